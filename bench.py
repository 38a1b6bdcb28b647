"""Benchmark: RCG iterations/s at 512^3, q = 4 (mixed precision with batched
line search vs the plain FP64 path) and FFT pass bandwidth, on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 512] [--q 4]

A step is one RCG iteration of maxdLqdt (GradL2ForLq, Hilbert gradient,
Leray projection, tangent projection, CG direction, mnbrak + brent line
search, vector transport, retraction) on the seeded divergence-free field
with spectrum exp(-|m|^2/256), B = 2 (synthetic data; the reference's input
fields are not available).  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_rcg(n, q, B, precision, batch, warmup, steps, seed=4, a=256.0, flags=0):
    """Run warmup+steps RCG iterations; time the last `steps` from the
    library's per-iteration records (each iteration ends in a synchronising
    reduction, so host stamps at the record are iteration boundaries)."""
    from maxlq_amd.lq import MaxLq
    import ctypes as C
    from maxlq_amd import capi

    g = MaxLq(n, q=q, B=B, precision=precision)
    h = g.field(3)
    g.seeded_field(h, seed, "gauss", a)
    g.lib.mlq_sync(g.ctx)
    stamps = {}
    counts = {}
    nJ = {}
    cur = {"it": 0, "nJ": 0}

    def cb(user, kind, it, i0, i1, d):
        if kind == capi.MLQ_LOG_ITER:
            stamps[i0] = time.perf_counter()
            counts[i0] = g.transform_counts()
            log(f"{precision} iter {i0} J={d[2]:.10e} tau={d[0]:.4e} beta={d[1]:.4e}")
        elif kind == capi.MLQ_LOG_BRENT and i0 == 999:
            pass

    cfg = capi.RcgCfg(warmup + steps, 0.0, 0, int(batch), 25, 10, int(flags))
    res = capi.RcgResult()
    t0 = time.perf_counter()
    rc = g.lib.mlq_rcg_solve(g.ctx, h, C.byref(cfg), C.byref(res), capi.LOG_CB(cb), None)
    g._chk(rc)
    t1 = time.perf_counter()
    if res.iters < warmup + steps:
        raise RuntimeError(f"{precision}: solve stopped after {res.iters} iterations "
                           f"(reason {res.stop_reason})")
    t_it = (stamps[warmup + steps] - stamps[warmup]) / steps
    g.tr_per_iter = tuple((counts[warmup + steps][k] - counts[warmup][k]) / steps for k in (0, 1))
    return g, h, res, t_it, t1 - t0


def iteration_algorithmic_bytes(g, h, batch=3):
    """Algorithmic bytes of one RCG iteration by kernel class: the library's
    per-launch count (one read of every input element a kernel needs, one
    write of every output element, pruned passes counted on their band),
    summed over the launches between two iteration records of a 2-iteration
    solve continuing from the field in h (profiling on, so not timed)."""
    import ctypes as C
    from maxlq_amd import capi
    snaps = []

    def cb(user, kind, it, i0, i1, d):
        if kind == capi.MLQ_LOG_ITER:
            snaps.append(g.prof())

    g.prof_enable(True)
    cfg = capi.RcgCfg(2, 0.0, 0, int(batch), 25, 10, 0)
    res = capi.RcgResult()
    g._chk(g.lib.mlq_rcg_solve(g.ctx, h, C.byref(cfg), C.byref(res), capi.LOG_CB(cb), None))
    g.prof_enable(False)
    if len(snaps) < 3:
        return None, None
    a, b = snaps[-2], snaps[-1]
    by = {k: b[k][2] - a[k][2] for k in b if b[k][2] - a[k][2] > 0}
    return sum(by.values()), by


def fft_roofline(g, batch=12, reps=5):
    """Per-pass average launch durations (HIP events on the library stream)
    and algorithmic bytes (input + output buffer bytes per launch)."""
    t, batch = g.time_fft(batch, reps, with_batch=True)   # the batch the scratch holds
    rb, sb = g.sizes()
    passes = ["x_fwd", "y_fwd", "z_fwd", "z_inv", "y_inv", "x_inv"]
    byts = [batch * (rb + sb), 2 * batch * sb, 2 * batch * sb,
            2 * batch * sb, 2 * batch * sb, batch * (rb + sb)]
    out = {}
    for p, ms, by in zip(passes, t, byts):
        out[p] = {"ms": ms, "bytes": by, "GBs": by / (ms * 1e-3) / 1e9}
    tot_b = sum(byts)
    tot_t = sum(t) * 1e-3
    return out, tot_b / tot_t / 1e9, (batch * rb + 2 * batch * sb) / (0.5 * tot_t) / 1e9


def pmc_traffic(kernel, n, precision):
    """HBM bytes per launch of one pass kernel (batch 12), from the committed
    summary of the FETCH_SIZE / WRITE_SIZE counter runs of this benchmark's
    transform sweep (counters scaled by the copy-kernel calibration of the
    same run).  None when no summary for this size and precision is present."""
    tag = "fp32" if precision == "mixed" else "fp64"
    path = os.path.join(ROOT, "profiles", f"pmc_fft_passes_{n}_{tag}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch:" and parts[1] == kernel:
            return float(parts[2])
    return None


def strided_tiles(n, wbytes):
    """(elements per thread, columns per workgroup) of the strided-pass tile
    shapes 0-3 (the library's tile table; shapes that do not fit fall back to
    shape 0)."""
    cmin = 16 if wbytes == 4 else 8
    e0 = 16 if n >= 1024 else 8
    c0 = max(cmin, 256 // (n // e0))
    out = []
    for v in range(4):
        er = e0 if v == 0 else (8 if v == 3 else 16)
        cr = c0 if v == 0 else (cmin if v == 1 else (2 * cmin if v == 2 else cmin // 2))
        ok = v == 0 or (n >= 128 and er <= n and (n // er) * cr <= 1024 and
                        n * cr * 2 * wbytes <= 160 * 1024)
        out.append((er, cr) if ok else (e0, c0))
    return out


def column_copy_rates(g, n, wbytes, nfields=12, reps=5):
    """GB/s of a 12-spectrum copy in the y and z access patterns of the
    strided passes, per tile shape (the better of the two workgroup orders)."""
    out = {}
    for axis, name in ((1, "y"), (2, "z")):
        rates = {}
        for v, (e, c) in enumerate(strided_tiles(n, wbytes)):
            best = 0.0
            for order in (0, 4):
                ms, by = g.time_column_copy(axis, v + order, nfields, reps)
                best = max(best, by / (ms * 1e-3) / 1e9)
            rates[f"E{e}xC{c}"] = round(best, 1)
        out[name] = rates
    return out


def iteration_traffic(n):
    """HBM bytes of one mixed RCG iteration (all kernels, FETCH_SIZE +
    WRITE_SIZE scaled by the copy calibration) from the committed summary of
    the counter runs of this benchmark; None when absent."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_rcg_iteration_{n}_q4_mixed.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if parts[:3] == ["per", "iteration:", "read"] and "total" in parts:
            return float(parts[parts.index("total") + 1]) * 1e9
    return None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads():
    """The host threads this process may use: its CPU affinity, capped by
    OMP_NUM_THREADS when that is set (the machine's share for this job)."""
    n = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return n


def cpu_baseline(n, q, B, threads, iters=2):
    """The reference-structured FP64 RCG on the host (the oracle's maxdLqdt,
    optimization.f90:7-514: gradient, Hilbert gradient, Leray projection,
    normal, CG direction, mnbrak + brent with its own J evaluations, vector
    transport, retraction), `iters` iterations on the bench's field at n^3
    (same seed, envelope scaled to the grid); the last iteration is timed,
    as the GPU line times iterations after a warm-up one."""
    from oracle import lq_oracle as L
    from oracle.seeded import seeded_field
    L.set_threads(threads)
    u = seeded_field(n, 4, "gauss", 256.0 * (n / 512.0) ** 2, q=q, B=B)
    orc = L.Oracle(L.Params(n=n, q=q, B=B, MAX_ITER=iters, OPTIM_TOL=0.0))
    stamps = []
    t0 = time.perf_counter()
    res = L.maxdLqdt(orc, u, on_start=lambda: stamps.append(time.perf_counter()),
                     on_iter=lambda rec, uu: stamps.append(time.perf_counter()))
    total = time.perf_counter() - t0
    if len(stamps) != iters + 1:
        raise RuntimeError(f"oracle solve stopped early ({res.stop_reason})")
    return stamps[-1] - stamps[-2], res.records[-1].nJ, total


def bench_linesearch8(args):
    """8 trial step sizes tau_i = 10^(-4 + 3i/7) at 512^3, q = 4, mixed: one
    batched launch vs 8 one-at-a-time launches after the same coefficient
    precompute, and the reference-structured evaluation (94 transforms per J)."""
    from maxlq_amd.lq import MaxLq
    n, q, B = args.n, args.q, args.B
    g = MaxLq(n, q=q, B=B, precision="mixed")
    hu, hd = g.field(3), g.field(3)
    g.seeded_field(hu, 4, "gauss", 256.0)
    g.seeded_field(hd, 14, "gauss", 256.0)
    taus = 10.0 ** (-4.0 + 3.0 * np.arange(8) / 7.0)
    g.time_linesearch(hu, hd, taus, 8)                      # warm-up
    Jb, tp, tb = g.time_linesearch(hu, hd, taus, 8)
    Js, tp1, ts = g.time_linesearch(hu, hd, taus, 1)
    Jr, _, tr = g.time_linesearch(hu, hd, taus, 0)
    Jp, tpp, tpe = g.time_linesearch(hu, hd, taus, -1)
    used, peak, _, _ = g.mem_info()
    out = {"metric": "8-step line search at 512^3 q=4 mixed: batched vs sequential (ms)",
           "value": tb, "unit": "ms", "higher_is_better": False,
           "config": {"workload": f"J_LineMin at 8 step sizes, {n}^3 q={q:g} B={B:g}, u seed 4, "
                                  "d seed 14, exp(-|m|^2/256)",
                      "taus": list(taus), "precompute_ms": tp, "batched_ms": tb,
                      "sequential_ms": ts, "reference_structured_ms": tr,
                      "sequential_over_batched": ts / tb, "reference_over_batched": tr / tb,
                      "polynomial_precompute_ms": tpp, "polynomial_eval_ms": tpe,
                      "max_rel_dJ_polynomial_vs_batched": float(np.max(np.abs(Jp - Jb) / np.abs(Jb))),
                      "max_rel_dJ_batched_vs_sequential": float(np.max(np.abs(Jb - Js) / np.abs(Js))),
                      "max_rel_dJ_batched_vs_reference": float(np.max(np.abs(Jb - Jr) / np.abs(Jr))),
                      "mem_peak_GB": peak / 1e9}}
    print(json.dumps(out), flush=True)


def bench_spectrum(args):
    """calculate_spectral_data (function_ops.f90:2614-2677) of a 512^3 field
    (seed 4, exp(-|m|^2/256)) on the GPU, FP64 and mixed, against the oracle
    (the reference's shell loop: n passes over the n^3 modes) on a 128^3 field,
    scaled by the loop's n^4 cost."""
    from maxlq_amd.lq import MaxLq
    n, q, B = args.n, args.q, args.B
    res = {}
    for prec in ("fp64", "mixed"):
        g = MaxLq(n, q=q, B=B, precision=prec)
        h = g.field(3)
        g.seeded_field(h, 4, "gauss", 256.0)
        g.spectral_data_field(h, True)                       # warm-up (allocations)
        ts = []
        for _ in range(5):
            t0 = time.perf_counter()
            sp = g.spectral_data_field(h, True)
            ts.append(time.perf_counter() - t0)
        w = 8 if prec == "fp64" else 4
        rb, sb = g.sizes()
        # 3 R2C transforms (field + half spectrum each), the shell pass over
        # 3 half spectra and the L2 norm's read of the field
        byts = 3 * (rb + sb) + 3 * sb + 3 * rb
        t = min(ts)
        res[prec] = {"ms": t * 1e3, "GBs": byts / t / 1e9, "bytes": byts,
                     "energy_check": float(2.0 * np.pi * sp[:, 1].sum())}
        g.close()
    from oracle.lq_oracle import Oracle, Params
    from oracle.seeded import seeded_field
    nc = 128
    u = seeded_field(nc, 4, "gauss", 256.0 * (nc / n) ** 2, q, B)
    orc = Oracle(Params(n=nc, q=q))
    t0 = time.perf_counter()
    orc.calculate_spectral_data(u, True)
    tc = time.perf_counter() - t0
    scale = (n / nc) ** 4
    out = {"metric": "energy spectrum in shells (calculate_spectral_data) at 512^3: GPU ms",
           "value": res["mixed"]["ms"], "unit": "ms", "higher_is_better": False,
           "config": {"workload": f"calculate_spectral_data of a {n}^3 field, seed 4, exp(-|m|^2/256), "
                                  "normalised by the L2 norm", "gpu": res,
                      "cpu_oracle": {"n": nc, "s": tc, "scaled_to_n_s": tc * scale,
                                     "scale": f"(n/{nc})^4, the reference's n shell passes over n^3 modes"}}}
    print(json.dumps(out), flush=True)


def bench_continuation(args):
    """256^3, q = 6, seed 3, |m|^2 exp(-|m|^2/64): B_j = 4^(j/9), j = 0..9,
    30 RCG iterations per B (OPTIM_TOL 0), mixed and FP64 each continuing
    from its own result; R error of the mixed path per step."""
    import ctypes as C
    from maxlq_amd import capi
    from maxlq_amd.lq import MaxLq
    n, q = 256, 6.0
    Bs = [4.0 ** (j / 9.0) for j in range(10)]
    ctx = {}
    for prec in ("mixed", "fp64"):
        g = MaxLq(n, q=q, B=Bs[0], precision=prec)
        h = g.field(3)
        g.seeded_field(h, 3, "k2gauss", 64.0)
        ctx[prec] = (g, h)
    steps = []
    for j, B in enumerate(Bs):
        row = {"B": B}
        for prec in ("mixed", "fp64"):
            g, h = ctx[prec]
            g.set_params(B=B)
            cfg = capi.RcgCfg(30, 0.0, 0, 3 if prec == "mixed" else 0, 25, 10, 0)
            res = capi.RcgResult()
            t0 = time.perf_counter()
            g._chk(g.lib.mlq_rcg_solve(g.ctx, h, C.byref(cfg), C.byref(res),
                                       C.cast(None, capi.LOG_CB), None))
            row[prec] = {"J": res.J1, "iters": res.iters, "s": time.perf_counter() - t0,
                         "n_J": res.n_J, "stop": res.stop_reason}
        g64, h64 = ctx["fp64"]
        gm, hm = ctx["mixed"]
        um = gm.download(hm)
        Rm = gm.calc_local_dLqdt_inclParts(um)[0]
        R64 = g64.calc_local_dLqdt_inclParts(um)[0]
        row["dR_over_R_same_field"] = abs(Rm - R64) / abs(R64)
        row["dJ_over_J_trajectories"] = abs(row["mixed"]["J"] - row["fp64"]["J"]) / abs(row["fp64"]["J"])
        log(f"B={B:.4f} J mixed {row['mixed']['J']:.10e} fp64 {row['fp64']['J']:.10e} "
            f"same-field dR/R {row['dR_over_R_same_field']:.2e}")
        steps.append(row)
    tm = sum(r["mixed"]["s"] for r in steps)
    t64 = sum(r["fp64"]["s"] for r in steps)
    im = sum(r["mixed"]["iters"] for r in steps)
    i64 = sum(r["fp64"]["iters"] for r in steps)
    out = {"metric": "B-continuation 256^3 q=6, 10 B x 30 RCG iterations: mixed vs FP64",
           "value": im / tm, "unit": "iter/s", "higher_is_better": True,
           "config": {"workload": "256^3 q=6 seed 3 |m|^2 exp(-|m|^2/64), B_j = 4^(j/9)",
                      "mixed_s": tm, "fp64_s": t64, "mixed_iters": im, "fp64_iters": i64,
                      "fp64_iter_per_s": i64 / t64,
                      "max_dR_over_R_same_field": max(r["dR_over_R_same_field"] for r in steps),
                      "steps": steps}}
    print(json.dumps(out), flush=True)


def bench_n1024(args):
    """1024^3, q = 9, seed 5, exp(-|m|^2/1024), B = 2.2, mixed: one full RCG
    iteration; allocator high-water mark and device memory against 288 GB."""
    from maxlq_amd.lq import MaxLq
    n, q, B = 1024, 9.0, 2.2
    probe = MaxLq(n, q=q, B=B, precision="mixed")
    path = {1: "resident-state", 2: "reference-structured, line search from coefficient fields",
            0: "reference-structured"}[probe.rcg_path()]
    probe.close()
    log(f"1024^3 q=9: {path} iteration")
    g, h, res, t_it, _ = run_rcg(n, q, B, "mixed", 3, 0, 1, seed=5, a=1024.0)
    used, peak, free, total = g.mem_info()
    out = {"metric": "RCG iteration 1024^3 q=9 mixed: s/iter and memory peak",
           "value": t_it, "unit": "s/iter", "higher_is_better": False,
           "config": {"workload": "1024^3 q=9 B=2.2 seed 5 exp(-|m|^2/1024)", "J1": res.J1,
                      "iteration": path,
                      "n_J": res.n_J, "mem_peak_GB": peak / 1e9, "mem_peak_GiB": peak / 2**30,
                      "mem_in_use_GB": used / 1e9, "device_total_GiB": total / 2**30,
                      "device_free_GB": free / 1e9}}
    print(json.dumps(out), flush=True)


def capi_flag_reference():
    from maxlq_amd import capi
    return capi.MLQ_RCG_REFERENCE


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--q", type=float, default=4.0)
    ap.add_argument("--B", type=float, default=2.0)
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=96,
                    help="grid of the timed host oracle iteration (cpu_baseline)")
    ap.add_argument("--fft-only", choices=["mixed", "fp64"], default=None,
                    help="only the batched transform sweep and a copy kernel (for profilers)")
    ap.add_argument("--case", choices=["iteration", "linesearch8", "continuation", "n1024",
                                       "spectrum"],
                    default="iteration",
                    help="iteration: the headline line (default); linesearch8: 8 step sizes "
                         "batched vs sequential at 512^3; continuation: 256^3 q=6 B-continuation "
                         "mixed vs FP64; n1024: one 1024^3 q=9 iteration and its memory peak; "
                         "spectrum: the shell energy spectrum at 512^3")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run it with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    n, q, B = args.n, args.q, args.B
    if args.case != "iteration":
        {"linesearch8": bench_linesearch8, "continuation": bench_continuation,
         "spectrum": bench_spectrum,
         "n1024": bench_n1024}[args.case](args)
        return
    if args.fft_only:
        from maxlq_amd.lq import MaxLq
        g = MaxLq(n, q=q, B=B, precision=args.fft_only)
        passes, gbs, strict = fft_roofline(g)
        rb, _ = g.sizes()
        tc = g.time_copy(12, 5)
        wb = 4 if args.fft_only == "mixed" else 8
        print(json.dumps({"fft_passes": passes, "fft_pass_GBs": gbs, "fft_strict_GBs": strict,
                          "copy": {"fields": 12, "ms": tc, "bytes": 2 * 12 * rb,
                                   "GBs": 2 * 12 * rb / (tc * 1e-3) / 1e9},
                          "column_copy_12_spectra_GBs": column_copy_rates(g, n, wb)}), flush=True)
        return

    log(f"mixed precision, batched line search: {n}^3 q={q} B={B}")
    gm, hm, resm, t_mixed, _ = run_rcg(n, q, B, "mixed", 3, args.warmup, args.steps)
    nJm = resm.n_J / max(resm.iters, 1)
    log(f"mixed: {resm.n_J} J requested, {resm.n_J_computed} computed, {resm.n_J_poly} by the "
        f"line-search polynomials")
    alg_bytes, alg_by_class = iteration_algorithmic_bytes(gm, hm)
    log(f"algorithmic bytes per iteration: {alg_bytes / 1e9:.1f} GB")
    mem_used, mem_peak, _, _ = gm.mem_info()
    fft_passes, fft_gbs, fft_strict = fft_roofline(gm)
    log(f"fft passes: " + ", ".join(f"{k}={v['GBs']:.0f}GB/s" for k, v in fft_passes.items()))
    # the practical ceiling of the same sweep: a plain copy of 12 fields
    rbm, _ = gm.sizes()
    tcopy = gm.time_copy(12, 5)
    copy_gbs = 2 * 12 * rbm / (tcopy * 1e-3) / 1e9
    colcopy = column_copy_rates(gm, n, 4)
    log("column copies (GB/s): " + json.dumps(colcopy))

    t_fp64 = None
    t_ref = None
    dR = None
    if not args.skip_fp64:
        log("FP64, reference-structured iteration (every operator recomputed from the field)")
        # one warm-up iteration: the first one allocates its temporaries
        gr, hr, _, t_ref, _ = run_rcg(n, q, B, "fp64", 0, 1, 1, flags=capi_flag_reference())
        gr.close()
        log("plain FP64 path")
        g64, h64, res64, t_fp64, _ = run_rcg(n, q, B, "fp64", 0, args.warmup, args.steps)
        # |dR|/R: R of the same field in both precisions (mixed iterate)
        u = gm.download(hm)
        Rm = gm.calc_local_dLqdt_inclParts(u)[0]
        R64 = g64.calc_local_dLqdt_inclParts(u)[0]
        dR = abs(Rm - R64) / abs(R64)
        del u
        g64.close()

    # dominant pass kernel for the roofline block
    dom = max(fft_passes, key=lambda k: fft_passes[k]["ms"])
    dp = fft_passes[dom]
    roof = {"bound": "hbm", "achieved": round(dp["GBs"], 1), "peak": PEAK_HBM_GBS,
            "unit": "GB/s", "frac": round(dp["GBs"] / PEAK_HBM_GBS, 4),
            "traffic": pmc_traffic(f"fft_{dom}", n, "mixed"),
            "kernel": f"fft_{dom}", "algorithmic_bytes": dp["bytes"], "launch_ms": round(dp["ms"], 4)}
    # the ceiling the pass's memory access pattern sets: the best copy rate of
    # that pattern with no arithmetic (any tile; x passes: the plain copy)
    axis = {"y": "y", "z": "z"}.get(dom[0])
    ceil = max(colcopy[axis].values()) if axis else copy_gbs
    roof["pattern_ceiling_GBs"] = round(ceil, 1)
    roof["frac_of_pattern_ceiling"] = round(dp["GBs"] / ceil, 4)

    cpu = None
    if not args.skip_cpu:
        thr = host_threads()
        nc = args.cpu_n
        log(f"cpu baseline: oracle RCG at {nc}^3 with {thr} threads")
        t_cpu, nJ_cpu, t_total = cpu_baseline(nc, q, B, thr)
        # n^3 log n work scaling of the transforms and pointwise passes
        scale = (n / nc) ** 3 * math.log(n) / math.log(nc)
        cpu = {"value": 1.0 / t_cpu, "unit": f"iter/s at {nc}^3", "cores": thr, "kind": "port",
               "sample": f"oracle maxdLqdt (numpy FP64, reference-structured) on the bench field "
                         f"at {nc}^3: 2 iterations in {t_total:.1f} s, the second timed "
                         f"({t_cpu:.2f} s, its own line search: n_J = {nJ_cpu}); transforms on "
                         f"{thr} threads (scipy.fft workers), pointwise work on one",
               "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
               "extrapolated_iter_per_s_at_bench_n": 1.0 / (t_cpu * scale),
               "extrapolation": f"x(n/{nc})^3 log(n)/log({nc}) = x{scale:.0f} to {n}^3"}

    out = {
        "metric": "RCG iter/s at 512^3 q=4: mixed vs FP64 (|dR|/R<1e-7); FFT HBM GB/s",
        "value": 1.0 / t_mixed,
        "unit": "iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_mixed * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64",
        "data": "synthetic",
        "config": {"workload": f"RCG iteration {n}^3 q={q:g} B={B:g} seeded exp(-|m|^2/256)",
                   "n": n, "q": q, "B": B, "seed": 4,
                   "fp64_iter_per_s": (1.0 / t_fp64) if t_fp64 else None,
                   "mixed_over_fp64": (t_fp64 / t_mixed) if t_fp64 else None,
                   "fp64_reference_structured_iter_per_s": (1.0 / t_ref) if t_ref else None,
                   "mixed_over_fp64_reference_structured": (t_ref / t_mixed) if t_ref else None,
                   "dR_over_R": dR, "n_J_per_iter_mixed": nJm, "mem_peak_GB_mixed": mem_peak / 1e9,
                   "n_J_poly_fraction_mixed": resm.n_J_poly / max(resm.n_J_computed, 1),
                   "fft_pass_GBs": round(fft_gbs, 1), "fft_strict_GBs": round(fft_strict, 1),
                   "fft_passes": {k: round(v["GBs"], 1) for k, v in fft_passes.items()},
                   "copy_12_fields_GBs": round(copy_gbs, 1),
                   "column_copy_12_spectra_GBs": colcopy,
                   "iteration_hbm_bytes": iteration_traffic(n),
                   "iteration_algorithmic_bytes": alg_bytes,
                   "iteration_algorithmic_bytes_by_class": alg_by_class,
                   "iteration_hbm_over_algorithmic": (round(iteration_traffic(n) / alg_bytes, 3)
                                                      if iteration_traffic(n) and alg_bytes
                                                      else None),
                   "iteration_algorithmic_GBs": (round(alg_bytes / t_mixed / 1e9, 1)
                                                 if alg_bytes else None),
                   "iteration_hbm_GBs": (round(iteration_traffic(n) / t_mixed / 1e9, 1)
                                         if iteration_traffic(n) else None),
                   "transforms_per_iter_mixed": {"forward": gm.tr_per_iter[0],
                                                 "inverse": gm.tr_per_iter[1],
                                                 "reference_c2c": 548 + 94 * nJm}},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
