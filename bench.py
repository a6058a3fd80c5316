#!/usr/bin/env python3
"""Headline benchmark: ABMC IC(0)-CG time to ||r||/||b|| < 1e-8 on one MI355X.

Workload (the reference's abmc_crs path at the configuration its metric is
quoted on): 3D 7-point variable-coefficient diffusion on a 192^3 Dirichlet
grid (n = 7,077,888), face coefficients 10^U[-2,2], diagonal = sum of the
cell's face coefficients + 1e-6, b ~ N(0,1), x0 = 0; ABMC ordering with
blocks of 64 rows (4x4x4 boxes, two colors; --block 64 takes breadth-first
blocks with first-fit greedy block coloring instead); the permuted system
P^T A P is solved as abmc_crs does (cg_crs.cpp:413-508).

A step is one complete solve to tolerance with every input resident in
HBM (the CG phase, as the reference times it, cg_crs.cpp:507-510).  The
headline runs the mixed-precision preconditioner (FP32 triangular solves,
FP64 CG recurrence, stop on the FP64 true residual); the plain FP64 path
runs beside it.  Prints ONE JSON line on stdout; progress goes to stderr.
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

METRIC = ("ABMC-ICCG time-to-1e-8 rel.res & HBM GB/s; served solves/s, p50/p99 latency")
HBM_PEAK_GBS = 8000.0
SEED = 0x1CC6


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def spmv_bytes(n, nnz, s=8, v=8):
    return nnz * (s + 4) + 4 * (n + 1) + 2 * n * v


def trsv_bytes(n, nnz_l, s, v_in, v):
    return nnz_l * (s + 4) + 4 * (n + 1) + n * v_in + 2 * n * v


def iter_bytes(n, nnz, nnz_l, mixed):
    if mixed:
        return (spmv_bytes(n, nnz) + trsv_bytes(n, nnz_l, 4, 8, 4) +
                trsv_bytes(n, nnz_l, 4, 4, 4) + 80 * n)
    return spmv_bytes(n, nnz) + 2 * trsv_bytes(n, nnz_l, 8, 8, 8) + 80 * n


def build_problem(N, block, dims=3):
    import iccg_amd as I
    t0 = time.time()
    nz = N if dims == 3 else 1
    A = I.gen_grid(N, N, nz, coef_exp=(-2.0, 2.0), diag_shift=1e-6, seed=SEED)
    log(f"grid {N}^{dims}: n={A.n} nnz={A.nnz} ({time.time() - t0:.1f}s)")
    t0 = time.time()
    if block == "box":
        # 64-row boxes: 4x4x4 in 3D, 8x8 in 2D
        bo, nb, bc, nc = (I.abmc_blocking_grid(N, N, N, 4, 4, 4) if dims == 3 else
                          I.abmc_blocking_grid(N, N, 1, 8, 8, 1))
    else:
        bo, nb, bc, nc = I.abmc_blocking(A, int(block))
    p = I.BuildPermutationByBlockColor(bo, bc)
    Ap = I.expand_lower_to_full(I.Permute_PtAP_LowerCRS_to_LowerCRS(A, p))
    cpb, boc, bp, rob = I.BuildABMCSchedule(p, bo, bc, nb, nc)
    sched = I.ColorSchedule.abmc(nc, nb, cpb, boc, bp, rob)
    log(f"ABMC blocks={nb} colors={nc} ({time.time() - t0:.1f}s)")
    b_old = I.gen_normal(A.n, SEED + 1)
    b = np.empty_like(b_old)
    b[p] = b_old  # b'[new_of_old[i]] = b[i]
    return Ap, sched, b, {"nb": nb, "nc": nc}


def time_solves(S, steps, warmup, max_iter, tol, true_res):
    """Timed solves (iterations replayed as hipGraphs), then one more solve
    launched kernel by kernel with HIP events on the solver's stream around
    every launch group, for the per-kernel durations."""
    S.set_timing(False)
    for _ in range(warmup):
        S.run(max_iter, tol, true_residual=true_res, zero_x0=True)
    res = []
    t0 = time.perf_counter()
    for _ in range(steps):
        res.append(S.run(max_iter, tol, true_residual=true_res, zero_x0=True))
    wall = time.perf_counter() - t0
    S.reset_timing()
    S.set_timing(True)
    prof = S.run(max_iter, tol, true_residual=true_res, zero_x0=True)
    S.set_timing(False)
    return res, wall, prof


def kernel_table(S, n, nnz, nnz_l, mixed, nc, iters):
    """Per kernel class, from HIP events recorded on the solver's stream
    around every launch group during the timed solves.  When the plan runs
    the last color forward and backward in one launch (abmc-sectioned-fused), the
    forward group holds that color's backward rows too, so the two sweeps
    are reported as one class."""
    import iccg_amd as I
    fwd = trsv_bytes(n, nnz_l, 4, 8, 4) if mixed else trsv_bytes(n, nnz_l, 8, 8, 8)
    bwd = trsv_bytes(n, nnz_l, 4, 4, 4) if mixed else trsv_bytes(n, nnz_l, 8, 8, 8)
    plan = S.M.plan()[0] if isinstance(S.M, I.IC0) else "none"
    fused = plan in ("abmc-sectioned-fused", "abmc-sectioned-fused-r")
    # with the r update and r.z in the sweeps, their bytes move from the
    # vector work (80n: r update 24n, x update 24n, r.z 8n with z in
    # registers, p update 24n) to the sweeps: the r update's 16n beyond the
    # sweep's own read of r, and r.z's read of r (8n); 48n stay
    ru = plan == "abmc-sectioned-fused-r"
    # multicolor: the r update's 16n into the forward sweep, r.z's read of r
    # (8n) into the backward one
    ru_mc = plan == "mc-sliced-r"
    cls = [("spmv_dots", [0], spmv_bytes(n, nnz), "one SpMV launch (k_spmv_cg)")]
    if fused:
        cls.append(("sweeps", [1, 2], fwd + bwd + (24 * n if ru else 0),
                    f"one forward and one backward sweep ({2 * nc - 1} color launches: "
                    "the last color forward and backward in one"
                    + ("; the sweeps also apply r -= alpha Ap and form r.z)" if ru else ")")))
    else:
        cls.append(("fwd_sweep", [1], fwd + (16 * n if ru_mc else 0),
                    f"one forward sweep ({nc} color launches"
                    + ("; with the r update r -= alpha Ap)" if ru_mc else ")")))
        cls.append(("bwd_sweep", [2], bwd + (8 * n if ru_mc else 0),
                    f"one backward sweep ({nc} color launches"
                    + ("; forming r.z)" if ru_mc else ")")))
    cls.append(("vector_ops", [3], 48 * n if ru or ru_mc else 80 * n,
                "one iteration's vector and scalar work (alpha finish, "
                + ("stop-test finish" if ru or ru_mc else "r update, r.z + stop-test finish")
                + ", x and p update)"))
    out = {}
    for nm, groups, byt, unit in cls:
        cnt, ms = S.timing(groups[0])
        if cnt == 0:
            continue
        ms = sum(S.timing(g)[1] for g in groups)
        per = cnt if groups[0] < 3 else max(iters, 1)
        avg = ms / per
        gbs = byt / (avg * 1e-3) / 1e9
        out[nm] = {"count": per, "avg_ms": round(avg, 5), "bytes": int(byt),
                   "gbs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4), "unit": unit}
    return out


_CORES = None  # physical cores, counted at start-up


def physical_cores():
    """Physical cores among the CPUs this process may run on: those CPUs
    divided by the hardware threads per core (the first CPU's sibling list),
    capped by OMP_NUM_THREADS when it is set.  Counted once, before OpenMP
    starts: with OMP_PROC_BIND set, OpenMP binds the main thread to one
    place, and the process's affinity shrinks to it."""
    global _CORES
    if _CORES is not None:
        return _CORES
    cpus = sorted(os.sched_getaffinity(0))
    tpc = 1
    try:
        with open(f"/sys/devices/system/cpu/cpu{cpus[0]}/topology/thread_siblings_list") as f:
            cnt = 0
            for part in f.read().strip().split(","):
                a, _, b = part.partition("-")
                cnt += int(b) - int(a) + 1 if b else 1
            tpc = max(1, cnt)
    except (OSError, ValueError):
        pass
    n = max(1, len(cpus) // tpc)
    cap = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    _CORES = min(n, cap) if cap > 0 else n
    return _CORES


def omp_protocol_env():
    """The reference's OpenMP placement for the CPU baseline (spread over
    physical cores), set before libgomp is loaded."""
    os.environ.setdefault("OMP_PROC_BIND", "spread")
    os.environ.setdefault("OMP_PLACES", "cores")


def cpu_baseline(Ap, sched, b, tol, max_iter, repeats, budget_s, label):
    """The reference's own abmc_crs CG (oracle/_ref, compiled from its
    sources with OpenMP) on the same permuted system and right-hand side,
    run to tolerance on the host's physical cores (OMP_PROC_BIND=spread,
    OMP_PLACES=cores), `repeats` times or until the budget is spent; the
    CG phase is timed as main() times it (cg_crs.cpp:507-510), the median
    reported.  Also returns the reference's x for the parity fields."""
    from oracle import ref as R
    from oracle.oracle import Csr, Schedule
    if not R.available():
        return None, None
    threads = physical_cores()
    R.set_threads(threads)
    A = Csr(Ap.n, Ap.rowptr, Ap.colind, Ap.val)
    if sched.mode == 1:
        s = Schedule(nc=sched.nc, mode=1, nb=sched.nb, color_ptr_blk=sched.color_ptr_blk,
                     blocks_of_color=sched.blocks_of_color, block_ptr=sched.block_ptr,
                     rows_of_block=sched.rows_of_block)
    else:
        s = Schedule(nc=sched.nc, color_ptr=sched.color_ptr, rows_of_color=sched.rows_of_color,
                     mode=0)
    t0 = time.time()
    R.ic0_factor_only(A, 0.0)
    log(f"reference IC(0) factorization (serial, untimed as in main()): {time.time() - t0:.1f}s")
    runs, x_ref = [], None
    t0 = time.time()
    for k in range(max(1, repeats)):
        it, rel, conv, x, sec = R.cg_timed_ic0(A, b, s, max_iter, tol=tol, build_factor=False)
        runs.append((sec, it, rel, conv))
        if x_ref is None:
            x_ref = x
        log(f"reference CG run {k + 1}: {it} iterations, rel {rel:.6e}, converged {conv}, "
            f"{sec:.3f}s on {threads} threads")
        if time.time() - t0 > budget_s:
            break
    secs = sorted(r[0] for r in runs)
    med = secs[len(secs) // 2] if len(secs) % 2 else 0.5 * (secs[len(secs) // 2 - 1] +
                                                             secs[len(secs) // 2])
    it, rel, conv = runs[0][1], runs[0][2], runs[0][3]
    out = {"value": round(med, 4), "unit": f"s (time to {tol:g}, CG phase)",
           "cores": threads, "kind": "reference",
           "sample": (f"the reference's {label} IC(0)-CG (oracle/_ref, OpenMP {threads} threads "
                      f"on {threads} physical cores, OMP_PROC_BIND="
                      f"{os.environ.get('OMP_PROC_BIND')}, OMP_PLACES="
                      f"{os.environ.get('OMP_PLACES')}) on the same permuted system and b, run "
                      f"to tolerance {len(runs)} times (median)"),
           "runs_s": [round(r[0], 4) for r in runs], "iters": it, "rel_resid": rel,
           "converged": conv}
    return out, x_ref


def refuse(msg):
    """A timed solve that misses its tolerance (or disagrees with the
    reference) gets no bench line."""
    print(f"bench.py: no result: {msg}", file=sys.stderr, flush=True)
    sys.exit(3)


def mc_system(A, seed_b):
    """Greedy (first-fit) multicolor ordering of A, the permuted system and
    its schedule, and b ~ N(0,1) in the permuted numbering, as picg_crs
    solves it (cg_crs.cpp:413-447)."""
    import iccg_amd as I
    color, nc = I.greedy_color(A)
    p = I.BuildPermutationByColor(color)
    Ap = I.expand_lower_to_full(I.Permute_PtAP_LowerCRS_to_LowerCRS(A, p))
    cp, rc = I.BuildColorPtrRows_FromPermutationAndColor(color, nc, p)
    sched = I.ColorSchedule.mc(nc, cp, rc)
    b_old = I.gen_normal(A.n, seed_b)
    b = np.empty_like(b_old)
    b[p] = b_old
    return Ap, sched, b


def batch_problem(count):
    """`count` 2D 5-point variable-coefficient systems: grid side uniform in
    [32, 256], face coefficients 10^U[-1,1] stored as FP32 values, b ~ N(0,1),
    greedy (red-black) multicolor ordering."""
    import iccg_amd as I
    rng = np.random.default_rng(SEED)
    sides = rng.integers(32, 257, size=count)
    out = []
    for s, N in enumerate(sides):
        A = I.gen_grid(int(N), int(N), 1, coef_exp=(-1.0, 1.0), seed=SEED + s, fp32_values=True)
        out.append(mc_system(A, SEED + 1_000_000 + s))
    return sides, out


def cpu_baseline_systems(systems, tol, budget_s, label):
    """The reference's picg_crs IC(0)-CG (oracle/_ref, OpenMP) on the first
    systems of the workload, one after another, until the budget is spent:
    solves per second of CG time (factorization untimed, as main() times
    it, cg_crs.cpp:507-510)."""
    from oracle import ref as R
    from oracle.oracle import Csr, Schedule
    if not R.available():
        return None
    threads = physical_cores()
    R.set_threads(threads)
    done, cg_s, t0 = 0, 0.0, time.time()
    for Ap, sched, b in systems:
        A = Csr(Ap.n, Ap.rowptr, Ap.colind, Ap.val)
        s = Schedule(nc=sched.nc, color_ptr=sched.color_ptr, rows_of_color=sched.rows_of_color,
                     mode=0)
        R.ic0_factor_only(A, 0.0)
        it, rel, conv, _, sec = R.cg_timed_ic0(A, b, s, 10000, tol=tol, build_factor=False)
        if not conv:
            refuse(f"the reference's CG did not converge on system {done}")
        cg_s += sec
        done += 1
        if time.time() - t0 > budget_s:
            break
    log(f"reference picg_crs: {done} solves, {cg_s:.2f}s of CG on {threads} threads")
    return {"value": round(done / cg_s, 3), "unit": "solves/s", "cores": threads,
            "kind": "reference",
            "sample": ((f"the first {done} systems of the {label}, solved one after another"
                        if len(systems) > 1 else f"the {label}, solved once") +
                       f" by the reference's picg_crs IC(0)-CG (oracle/_ref, OpenMP {threads} "
                       f"threads on {threads} physical cores, OMP_PROC_BIND="
                       f"{os.environ.get('OMP_PROC_BIND')}, OMP_PLACES="
                       f"{os.environ.get('OMP_PLACES')}), FP64, tol {tol:g}; CG phase timed")}


def batch_kernel_table(B, systems, res, count, max_iter, tol, M32, dA):
    """One more batched solve of every system with the batch's timing mode
    (kernels launched one by one, HIP events on the batch's stream around
    every launch group).  Per class: summed event time and the algorithmic
    bytes of every system over the iterations it ran (FP32 A values, FP32
    factor and y/z), hence GB/s over the whole batched solve."""
    for s in range(count):
        B.remove(s)
    B.set_timing(True)
    B.reset_timing()
    slots = [B.add(dA[k], M32[k], systems[k][2], max_iter=max_iter, tol=tol)
             for k in range(count)]
    while B.run():
        pass
    B.set_timing(False)
    it = [B.result(s)[1].iters for s in slots]
    fwd = bwd = spmv = vec = 0
    for (A, _, _), k in zip(systems, it):
        n, nnz = A.n, A.nnz
        nnz_l = (nnz + n) // 2
        spmv += k * spmv_bytes(n, nnz, s=4)
        fwd += k * trsv_bytes(n, nnz_l, 4, 8, 4)
        bwd += k * trsv_bytes(n, nnz_l, 4, 4, 4)
        vec += k * 80 * n
    out = {}
    for c, (nm, byt) in enumerate([("spmv_dots", spmv), ("fwd_sweep", fwd), ("bwd_sweep", bwd),
                                   ("vector_ops", vec)]):
        cnt, ms = B.timing(c)
        gbs = byt / (ms * 1e-3) / 1e9 if ms > 0 else None
        out[nm] = {"launch_groups": cnt, "total_ms": round(ms, 3), "bytes": int(byt),
                   "gbs": round(gbs, 1) if gbs else None,
                   "frac": round(gbs / HBM_PEAK_GBS, 4) if gbs else None}
    log("batched kernel classes: " + ", ".join(f"{k} {v['total_ms']} ms {v['gbs']} GB/s"
                                              for k, v in out.items()))
    return out


def batch_roofline(kernels):
    """The kernel class with the largest event time over the whole batched
    solve: its algorithmic bytes (every system's iterations) over that time."""
    nm, k = max(kernels.items(), key=lambda kv: kv[1]["total_ms"])
    return {"bound": "hbm", "kernel": nm,
            "unit_of_launch": "every launch of the class over one batched solve of all systems",
            "achieved": k["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": k["frac"],
            "traffic": None, "algorithmic_bytes": k["bytes"], "total_ms": k["total_ms"]}


def main_batch(args):
    """Many-system workload: one Batch solving every system (mixed precision,
    stop on the FP64 true residual) against the plain FP64 single-system
    solver run on the systems one at a time."""
    import iccg_amd as I
    count, tol, max_iter = args.systems, args.tol, args.max_iter
    t0 = time.time()
    sides, systems = batch_problem(count)
    n_tot = sum(A.n for A, _, _ in systems)
    nnz_tot = sum(A.nnz for A, _, _ in systems)
    log(f"{count} systems, sum n = {n_tot}, sum nnz = {nnz_tot} ({time.time() - t0:.1f}s)")
    t0 = time.time()
    dA = [I.DeviceCSR(A) for A, _, _ in systems]
    M32 = [I.IC0(d, 0.0, sched=s, precision="mixed") for d, (_, s, _) in zip(dA, systems)]
    log(f"device setup + IC(0) mixed: {time.time() - t0:.1f}s")

    B = I.Batch(count)

    def batch_step():
        slots = [B.add(dA[k], M32[k], systems[k][2], max_iter=max_iter, tol=tol)
                 for k in range(count)]
        while B.run():
            pass
        res = [B.result(s)[1] for s in slots]
        return slots, res

    for _ in range(args.warmup):
        slots, _ = batch_step()
        for s in slots:
            B.remove(s)
    t0 = time.perf_counter()
    for k in range(args.steps):
        slots, res = batch_step()
        if k < args.steps - 1:
            for s in slots:
                B.remove(s)
    wall = time.perf_counter() - t0
    for r in res:
        if not (r.converged and r.true_rel_resid < tol):
            refuse(f"a batched system missed tol: {r.iters} iterations, true "
                   f"{r.true_rel_resid:.3e}")
    xs = [B.x(s) for s in slots]
    chunk, cw = B.tuning
    kernels = batch_kernel_table(B, systems, res, count, max_iter, tol, M32, dA)
    sps = count * args.steps / wall
    iters = np.array([r.iters for r in res])
    trel = np.array([r.true_rel_resid for r in res])
    log(f"batched mixed: {sps:.1f} solves/s ({wall / args.steps * 1e3:.1f} ms per {count}); "
        f"iters {iters.min()}..{iters.max()} (mean {iters.mean():.1f}); max true rel "
        f"{trel.max():.2e}; all converged {all(r.converged for r in res)}; "
        f"tuning chunk={chunk} slices, sweep={cw} slices/wave")
    del B

    # one at a time: the plain FP64 single-system solver, then mixed
    def one_at_a_time(precision):
        Ms = [I.IC0(d, 0.0, sched=s, precision=precision) for d, (_, s, _) in zip(dA, systems)]
        Ss = []
        for d, M, (_, _, b) in zip(dA, Ms, systems):
            S = I.Solver(d, M)
            S.set(b)
            Ss.append(S)
        tr = precision == "mixed"
        for S in Ss:  # warm: graphs captured
            S.run(max_iter, tol, true_residual=tr, zero_x0=True)
        t0 = time.perf_counter()
        out = [S.run(max_iter, tol, true_residual=tr, zero_x0=True) for S in Ss]
        w = time.perf_counter() - t0
        xs1 = [S.x() for S in Ss]
        del Ss, Ms
        return count / w, out, xs1

    if args.no_compare:
        log("comparison legs skipped (--no-compare)")
        return
    sps64, res64, xs64 = one_at_a_time("fp64")
    if not all(r.converged for r in res64):
        refuse("a one-at-a-time FP64 solve did not converge")
    tr64 = np.array([r.true_rel_resid for r in res64])
    log(f"one-at-a-time FP64: {sps64:.1f} solves/s; max true rel {tr64.max():.2e}")
    sps32_1, res32_1, _ = one_at_a_time("mixed")
    log(f"one-at-a-time mixed: {sps32_1:.1f} solves/s")
    err = max(float(np.linalg.norm(a - c) / np.linalg.norm(c)) for a, c in zip(xs, xs64))

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_systems(systems, tol, args.cpu_budget,
                                   f"{count}-system workload")
    line = {
        "metric": METRIC, "value": round(sps, 2), "unit": "solves/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64 CG + f32 values and factors",
        "data": "synthetic",
        "config": {"workload": (f"batched ICCG: {count} independent 2D 5-pt variable-coefficient "
                                f"systems, grid side U[32,256] (sum n={n_tot}, sum nnz={nnz_tot}), "
                                f"coefficients 10^U[-1,1] stored as FP32, b~N(0,1), greedy "
                                f"multicolor, FP32 CRS values and factors + FP64 CG and dots, "
                                f"tol {tol:g} on the FP64 true residual; a step solves all "
                                f"{count}"),
                   "systems": count, "sum_n": n_tot, "sum_nnz": nnz_tot, "tol": tol},
        "one_at_a_time_fp64_solves_per_s": round(sps64, 2),
        "speedup_vs_one_at_a_time_fp64": round(sps / sps64, 3),
        "one_at_a_time_mixed_solves_per_s": round(sps32_1, 2),
        "iters_min_mean_max": [int(iters.min()), round(float(iters.mean()), 1), int(iters.max())],
        "max_true_rel_resid": float(trel.max()),
        "fp64_max_true_rel_resid": float(tr64.max()),
        "max_err_vs_fp64": err,
        "tuning": {"chunk_slices": chunk, "sweep_slices_per_wave": cw},
        "kernels": kernels,
        "roofline": batch_roofline(kernels),
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def serve_requests(count, seed):
    """The first `count` requests of the serving stream: 2D 5-pt or 3D 7-pt
    (50/50) variable-coefficient Laplacians, n log-uniform in [1e4, 4e6]
    (grid side rounded), face coefficients 10^U[-2,2], diagonal = sum of the
    face coefficients, 1..8 right-hand sides ~ N(0,1)."""
    import iccg_amd as I
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(count):
        d = 2 if rng.random() < 0.5 else 3
        n_target = 10.0 ** rng.uniform(4.0, np.log10(4e6))
        side = max(2, int(round(n_target ** (1.0 / d))))
        dims = (side, side, 1) if d == 2 else (side, side, side)
        nrhs = int(rng.integers(1, 9))
        A = I.gen_grid(*dims, coef_exp=(-2.0, 2.0), seed=seed + k)
        b = np.stack([I.gen_normal(A.n, seed * 1000 + 16 * k + j) for j in range(nrhs)])
        reqs.append((A, b, dims))
    return reqs


def drive(server, reqs, arrivals):
    """Submits every request at its arrival time, then collects all of them;
    returns per-request (latency_ms, prepare_ms, results, x) and the span
    from the first arrival to the last completion."""
    t0 = time.perf_counter()
    ids = []
    for (A, b, dims), ta in zip(reqs, arrivals):
        dt = ta - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        ids.append(server.submit(A, b, grid=dims))
    # collected as they finish (the progress lines are completion times),
    # returned in submission order
    where = {rid: k for k, rid in enumerate(ids)}
    out = [None] * len(ids)
    for done in range(1, len(ids) + 1):
        rid = server.wait()
        x, res, lat, prep = server.result(rid)
        out[where[rid]] = (lat, prep, res, x)
        server.release(rid)
        if done % 16 == 0 or done == len(ids):
            log(f"  {done}/{len(ids)} requests done ({time.perf_counter() - t0:.1f}s)")
    return out, time.perf_counter() - t0


def main_mc128(args):
    """picg_crs at the red-black configuration: 3D 7-point Laplacian 128^3
    (n = 2,097,152), greedy (red-black) multicolor IC(0), FP64, b ~ N(0,1),
    x0 = 0, tol 1e-8; time to tolerance with SpMV and sweep GB/s."""
    import iccg_amd as I
    N = 128
    t0 = time.time()
    A = I.gen_grid(N, N, N, seed=SEED)
    Ap, sched, b = mc_system(A, SEED + 1)
    n, nnz = Ap.n, Ap.nnz
    nnz_l = (nnz + n) // 2
    log(f"grid {N}^3 red-black: n={n} nnz={nnz} colors={sched.nc} ({time.time() - t0:.1f}s)")
    dA = I.DeviceCSR(Ap)
    M = I.IC0(dA, 0.0, sched=sched, precision="fp64")
    S = I.Solver(dA, M)
    S.set(b)
    res, wall, prof = time_solves(S, args.steps, args.warmup, args.max_iter, args.tol, False)
    ms = wall / args.steps * 1e3
    check_solves("fp64", res, args.tol, true_resid=True)
    r = res[-1]
    x = S.x()
    k = kernel_table(S, n, nnz, nnz_l, False, sched.nc, prof.iters)
    gbs = iter_bytes(n, nnz, nnz_l, False) * r.iters / (ms * 1e-3) / 1e9
    log(f"fp64: {r.iters} iters, rel {r.rel_resid:.3e}, true {r.true_rel_resid:.3e}, "
        f"{ms:.1f} ms/solve, {gbs:.0f} GB/s")
    cpu, parity = None, None
    if not args.no_cpu:
        cpu, x_ref = cpu_baseline(Ap, sched, b, args.tol, args.max_iter, 10, args.cpu_budget,
                                  "picg_crs")
        if cpu is not None:
            parity = reference_parity(cpu, x_ref, r, x, x)
            parity.pop("err_mixed_vs_reference")
    dom = max(k.items(), key=lambda kv: kv[1]["avg_ms"] * kv[1]["count"])
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_mc128_{dom[0]}.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_launch")
    line = {
        "metric": METRIC, "value": round(ms / 1e3, 4), "unit": "s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": (f"picg_crs: 3D 7-pt Laplacian {N}^3 (n={n}, nnz={nnz}), greedy "
                                f"red-black IC(0), FP64, b~N(0,1), x0=0, tol {args.tol:g}"),
                   "n": n, "nnz": nnz, "tol": args.tol},
        "iters": r.iters, "rel_resid": r.rel_resid, "true_rel_resid": r.true_rel_resid,
        "hbm_gbs": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4), "kernels": k,
        "roofline": {"bound": "hbm", "kernel": dom[0], "unit_of_launch": dom[1]["unit"],
                     "achieved": dom[1]["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": dom[1]["frac"], "traffic": traffic,
                     "algorithmic_bytes_per_launch": dom[1]["bytes"],
                     "avg_ms": dom[1]["avg_ms"]},
        "cpu_baseline": cpu, "reference_parity": parity,
    }
    print(json.dumps(line), flush=True)


def main_serve(args):
    """Serving workload: the same Poisson-arrival request stream through the
    batched request loop (mixed precision, continuous batching) and through
    the plain one-at-a-time FP64 path."""
    import iccg_amd as I
    count, tol = args.requests, args.tol
    t0 = time.time()
    reqs = serve_requests(count, SEED)
    rng = np.random.default_rng(SEED + 7)
    arrivals = np.cumsum(rng.exponential(1.0 / args.rate, size=count))
    arrivals -= arrivals[0]
    n_rhs = sum(b.shape[0] for _, b, _ in reqs)
    ns = np.array([A.n for A, _, _ in reqs])
    log(f"{count} requests ({n_rhs} right-hand sides), n {ns.min()}..{ns.max()} "
        f"(median {int(np.median(ns))}), arrivals over {arrivals[-1]:.2f}s "
        f"({time.time() - t0:.1f}s to generate)")

    def summary(out, span, label, mixed):
        lat = np.array([o[0] for o in out])
        prep = np.array([o[1] for o in out])
        trel = max(r.true_rel_resid for o in out for r in o[2])
        conv = all(r.converged for o in out for r in o[2])
        # algorithmic HBM bytes of every iteration of every right-hand side
        byt = sum(iter_bytes(A.n, A.nnz, (A.nnz + A.n) // 2, mixed) * r.iters
                  for (A, _, _), o in zip(reqs, out) for r in o[2])
        d = {"solves_per_s": round(n_rhs / span, 3), "requests_per_s": round(count / span, 3),
             "alg_bytes": int(byt), "alg_gbs": round(byt / span / 1e9, 1),
             "sum_rhs_solve_s": round(sum(r.solve_ms for o in out for r in o[2]) / 1e3, 3),
             "p50_latency_ms": round(float(np.percentile(lat, 50)), 2),
             "p99_latency_ms": round(float(np.percentile(lat, 99)), 2),
             "mean_prepare_ms": round(float(prep.mean()), 2), "span_s": round(span, 3),
             "max_true_rel_resid": trel, "all_converged": conv}
        log(f"{label}: {d}")
        return d

    served_mode = I.Server("batched", precision="mixed", max_slots=args.slots,
                           workers=args.workers, hbm_budget=int(args.hbm_budget * (1 << 30)),
                           multi_rows=args.multi_rows, multi_runners=args.multi_runners)
    out_b, span_b = drive(served_mode, reqs, arrivals)
    st = served_mode.stats()
    engine = {k: round(st[k], 1) for k in ("engine_run_ms", "engine_admit_ms",
                                           "engine_harvest_ms")}
    log(f"batched engine thread: {engine} (span {span_b * 1e3:.0f} ms)")
    del served_mode
    sb = summary(out_b, span_b, "batched mixed", True)
    sb["engine_ms"] = engine
    ss, err = None, None
    if not args.no_compare:
        seq = I.Server("sequential", precision="fp64")
        out_s, span_s = drive(seq, reqs, arrivals)
        del seq
        ss = summary(out_s, span_s, "one-at-a-time fp64", False)
        err = max(float(np.linalg.norm(a[3] - c[3]) / np.linalg.norm(c[3]))
                  for a, c in zip(out_b, out_s))
    line = {
        "metric": METRIC, "value": sb["solves_per_s"], "unit": "solves/s",
        "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": round(span_b * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64 CG + f32 preconditioner", "data": "synthetic",
        "config": {"workload": (f"serving loop: the first {count} requests of a Poisson stream "
                                f"(rate {args.rate:g}/s) of 2D/3D variable-coefficient "
                                f"Laplacians, n log-uniform [1e4,4e6], coefficients "
                                f"10^U[-2,2], 1-8 right-hand sides N(0,1) ({n_rhs} in all); "
                                f"ABMC (box blocks of 64 rows) IC(0)-CG to {tol:g}; a step "
                                f"serves the whole sample, latency = submit to result"),
                   "requests": count, "right_hand_sides": n_rhs, "rate_per_s": args.rate,
                   "max_slots": args.slots, "workers": args.workers,
                   "multi_rows": args.multi_rows, "multi_runners": args.multi_runners,
                   "hbm_budget_gib": args.hbm_budget},
        "p50_latency_ms": sb["p50_latency_ms"], "p99_latency_ms": sb["p99_latency_ms"],
        "batched_mixed": sb, "one_at_a_time_fp64": ss,
        "speedup_vs_one_at_a_time_fp64": (round(sb["solves_per_s"] / ss["solves_per_s"], 3)
                                          if ss else None),
        "max_err_vs_fp64": err,
    }
    print(json.dumps(line), flush=True)


def main_multi(args):
    """Several right-hand sides of one system: the multi-right-hand-side
    solver (each right-hand side its own CG recurrence, the operator read
    once per set of up to four) against the same right-hand sides solved one
    after another by the single-system solver, both mixed precision with the
    FP64 true-residual stop, on the headline system (or --dims 2)."""
    import iccg_amd as I
    N, K = args.size, args.nrhs
    Ap, sched, b, info = build_problem(N, args.block, args.dims)
    n, nnz = Ap.n, Ap.nnz
    B = np.stack([b] + [I.gen_normal(n, SEED + 100 + k) for k in range(1, K)])
    dA = I.DeviceCSR(Ap)
    M = I.IC0(dA, 0.0, sched=sched, precision=args.precision)
    true_res = args.precision == "mixed"
    MS = I.MultiSolver(dA, M, K)
    MS.set(B)
    for _ in range(args.warmup):
        MS.run(args.max_iter, args.tol, true_residual=true_res, zero_x0=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = MS.run(args.max_iter, args.tol, true_residual=true_res, zero_x0=True)
    wall_m = (time.perf_counter() - t0) / args.steps
    X = MS.x()
    for r in res:
        check_solves("multi", [r], args.tol, true_resid=True)
    log(f"multi ({K} right-hand sides, layout {MS.layout()}): {wall_m * 1e3:.1f} ms, "
        f"iters {[r.iters for r in res]}")
    # one more solve launched kernel by kernel, HIP events around each class
    MS.reset_timing()
    MS.set_timing(True)
    MS.run(args.max_iter, args.tol, true_residual=true_res, zero_x0=True)
    MS.set_timing(False)
    classes = {}
    for c, name in enumerate(["spmv", "fwd_sweep", "bwd_sweep", "vector_finish"]):
        cnt, ms = MS.timing(c)
        classes[name] = {"launch_groups": cnt, "avg_us": round(ms / max(cnt, 1) * 1e3, 2)}
    log(f"multi per set iteration (events): {classes}")
    del MS
    if args.no_compare:
        print(json.dumps({"metric": "right-hand sides of one system solved to tol per second",
                          "value": round(K / wall_m, 3), "unit": "right-hand sides/s",
                          "ms_per_step": round(wall_m * 1e3, 3), "nrhs": K,
                          "iters": [r.iters for r in res],
                          "classes_per_set_iteration": classes}), flush=True)
        return
    S = I.Solver(dA, M)
    single, xs, wall_s = [], [], 0.0
    for k in range(K):
        S.set(B[k])
        for _ in range(args.warmup):
            S.run(args.max_iter, args.tol, true_residual=true_res, zero_x0=True)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            r = S.run(args.max_iter, args.tol, true_residual=true_res, zero_x0=True)
        wall_s += (time.perf_counter() - t0) / args.steps
        single.append(r)
        xs.append(S.x())
    same = all((a.iters, a.rel_resid, a.restarts) == (c.iters, c.rel_resid, c.restarts)
               for a, c in zip(res, single)) and np.array_equal(X, np.stack(xs))
    if not same:
        refuse("the multi-right-hand-side solves differ from the single-system ones")
    log(f"one at a time: {wall_s * 1e3:.1f} ms for {K} right-hand sides "
        f"({wall_m and wall_s / wall_m:.2f}x); results bitwise equal")
    line = {
        "metric": "right-hand sides of one system solved to tol per second",
        "value": round(K / wall_m, 3), "unit": "right-hand sides/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall_m * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": ("f64 CG + f32 preconditioner" if true_res else "f64"),
        "data": "synthetic",
        "config": {"workload": (f"{K} right-hand sides of {args.dims}D "
                                f"{2 * args.dims + 1}-pt variable-coefficient diffusion "
                                f"{N}^{args.dims} (n={n}, nnz={nnz}), ABMC B={args.block} "
                                f"({info['nb']} blocks, {info['nc']} colors), b~N(0,1), "
                                f"tol {args.tol:g}"), "n": n, "nnz": nnz, "nrhs": K},
        "iters": [r.iters for r in res], "classes_per_set_iteration": classes,
        "one_at_a_time": {"value": round(K / wall_s, 3), "ms": round(wall_s * 1e3, 3)},
        "speedup_vs_one_at_a_time": round(wall_s / wall_m, 3),
        "bitwise_equal_to_one_at_a_time": bool(same),
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["abmc192", "mc128", "batch", "serve", "multi"],
                    default="abmc192",
                    help="abmc192: the single-system headline; mc128: red-black FP64 at 128^3; "
                         "batch: many systems at once; serve: the request loop; multi: several "
                         "right-hand sides of the abmc192 system (--nrhs)")
    ap.add_argument("--nrhs", type=int, default=4)
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed",
                    help="multi workload: the preconditioner's precision")
    ap.add_argument("--systems", type=int, default=512)
    ap.add_argument("--requests", type=int, default=64)
    ap.add_argument("--no-compare", action="store_true",
                    help="batch, serve and multi workloads: time the batched / served / "
                         "multi-right-hand-side solves only (for profiling)")
    ap.add_argument("--rate", type=float, default=20.0, help="request arrivals per second")
    ap.add_argument("--slots", type=int, default=256)
    ap.add_argument("--workers", type=int, default=12)
    ap.add_argument("--multi-rows", type=int, default=50000,
                    help="serve: requests of at least this many rows run on the "
                         "multi-right-hand-side solver beside the batch (0: none)")
    ap.add_argument("--multi-runners", type=int, default=6)
    ap.add_argument("--hbm-budget", type=float, default=96.0, help="GiB")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=192)
    ap.add_argument("--dims", type=int, choices=[2, 3], default=3,
                    help="3: the N^3 7-point headline; 2: an N^2 5-point grid of the same "
                         "coefficients (8x8 boxes), as the serving stream's 2D requests")
    ap.add_argument("--block", default="box",
                    help="'box': 4x4x4 grid boxes of 64 rows; or an integer B for breadth-first "
                         "blocks of B rows (both colored first-fit)")
    ap.add_argument("--max-iter", type=int, default=20000)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=30.0,
                    help="seconds of reference CG runs to tolerance (at least one runs)")
    ap.add_argument("--cpu-repeats", type=int, default=3)
    ap.add_argument("--skip-fp64", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the multi-GPU design "
              "(sharding by row slabs or by whole systems) is not run here", file=sys.stderr)
        sys.exit(2)

    physical_cores()  # before anything starts OpenMP
    if args.workload in ("abmc192", "mc128", "batch"):
        omp_protocol_env()
    import iccg_amd as I
    if I.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.workload == "batch":
        return main_batch(args)
    if args.workload == "serve":
        return main_serve(args)
    if args.workload == "mc128":
        return main_mc128(args)
    if args.workload == "multi":
        return main_multi(args)
    N = args.size
    Ap, sched, b, info = build_problem(N, args.block, args.dims)
    n, nnz = Ap.n, Ap.nnz
    nnz_l = (nnz + n) // 2

    t0 = time.time()
    dA = I.DeviceCSR(Ap)
    M32 = I.IC0(dA, 0.0, sched=sched, precision="mixed")
    log(f"device setup + IC(0) (mixed): {time.time() - t0:.1f}s")
    S32 = I.Solver(dA, M32)
    S32.set(b)
    res32, wall32, prof32 = time_solves(S32, args.steps, args.warmup, args.max_iter, args.tol,
                                        True)
    ms32 = wall32 / args.steps * 1e3
    check_solves("mixed", res32, args.tol, true_resid=True)
    r32 = res32[-1]
    x32 = S32.x()
    k32 = kernel_table(S32, n, nnz, nnz_l, True, info["nc"], prof32.iters)
    log(f"mixed: {r32.iters} iters, true rel {r32.true_rel_resid:.3e}, {ms32:.1f} ms/solve, "
        f"restarts {r32.restarts}")
    del S32

    fp64 = None
    if not args.skip_fp64:
        t0 = time.time()
        M64 = I.IC0(dA, 0.0, sched=sched, precision="fp64")
        log(f"IC(0) (fp64): {time.time() - t0:.1f}s")
        S64 = I.Solver(dA, M64)
        S64.set(b)
        res64, wall64, prof64 = time_solves(S64, args.steps, args.warmup, args.max_iter,
                                            args.tol, False)
        ms64 = wall64 / args.steps * 1e3
        check_solves("fp64", res64, args.tol, true_resid=True)
        r64 = res64[-1]
        x64 = S64.x()
        k64 = kernel_table(S64, n, nnz, nnz_l, False, info["nc"], prof64.iters)
        err = float(np.linalg.norm(x32 - x64) / np.linalg.norm(x64))
        log(f"fp64: {r64.iters} iters, rel {r64.rel_resid:.3e}, true {r64.true_rel_resid:.3e}, "
            f"{ms64:.1f} ms/solve; mixed vs fp64 rel err {err:.2e}")
        fp64 = {"time_s": round(ms64 / 1e3, 4), "iters": r64.iters,
                "rel_resid": r64.rel_resid, "true_rel_resid": r64.true_rel_resid,
                "gbs": round(iter_bytes(n, nnz, nnz_l, False) * r64.iters / (ms64 * 1e-3) / 1e9, 1),
                "kernels": k64}
        del S64, M64

    # practical bandwidth of a plain device copy (1 GiB read + 1 GiB write)
    copy_ms = I.bench_copy(1 << 30, 20)
    copy_gbs = round(2 * (1 << 30) / (copy_ms * 1e-3) / 1e9, 1)
    log(f"copy kernel: {copy_gbs} GB/s")

    bytes_it = iter_bytes(n, nnz, nnz_l, True)
    gbs32 = bytes_it * r32.iters / (ms32 * 1e-3) / 1e9
    # the dominant single kernel of the headline solve: the one launch with
    # the largest share of time (the sweeps class is three launches of
    # different kernels, each shorter than the SpMV)
    dom = ("spmv_dots", k32["spmv_dots"])
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{dom[0]}.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_launch")
    roofline = {"bound": "hbm", "kernel": dom[0], "unit_of_launch": dom[1]["unit"],
                "achieved": dom[1]["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": dom[1]["frac"], "traffic": traffic,
                "copy_kernel_gbs": copy_gbs,
                "algorithmic_bytes_per_launch": dom[1]["bytes"], "avg_ms": dom[1]["avg_ms"]}
    if dom[0] == "spmv_dots":
        # the kernel's own layout: 64-row slices padded to their widest row,
        # FP64 values with 16-bit columns, one 16-entry block table per slice
        w = np.diff(Ap.rowptr)
        pad = np.zeros(((n + 63) // 64) * 64, np.int64)
        pad[:n] = w
        padded = int(pad.reshape(-1, 64).max(axis=1).sum()) * 64
        ns = (n + 63) // 64
        lay = padded * (8 + 2) + ns * 64 + 4 * (ns + 1) + 2 * n * 8
        roofline["kernel_layout_bytes_per_launch"] = int(lay)
        roofline["kernel_layout_gbs"] = round(lay / (dom[1]["avg_ms"] * 1e-3) / 1e9, 1)

    cpu, parity = None, None
    if not args.no_cpu:
        cpu, x_ref = cpu_baseline(Ap, sched, b, args.tol, args.max_iter, args.cpu_repeats,
                                  args.cpu_budget, "abmc_crs")
        if cpu is not None:
            parity = reference_parity(cpu, x_ref, r64 if fp64 else None,
                                      x64 if fp64 else None, x32)

    line = {
        "metric": METRIC, "value": round(ms32 / 1e3, 4), "unit": "s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms32, 3), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64 CG + f32 preconditioner", "data": "synthetic",
        "config": {"workload": (f"abmc_crs: {args.dims}D {2 * args.dims + 1}-pt "
                                f"variable-coefficient diffusion {N}^{args.dims} "
                                f"(n={n}, nnz={nnz}), ABMC "
                                + (("4x4x4 boxes" if args.dims == 3 else "8x8 boxes")
                                   + " of 64 rows" if args.block == "box" else
                                   f"breadth-first blocks of {args.block} rows")
                                + f" ({info['nb']} blocks, {info['nc']} colors by first-fit "
                                f"greedy coloring), b~N(0,1), x0=0, "
                                f"tol {args.tol:g} on the FP64 true residual"),
                   "n": n, "nnz": nnz, "tol": args.tol},
        "iters": r32.iters, "true_rel_resid": r32.true_rel_resid, "restarts": r32.restarts,
        "hbm_gbs": round(gbs32, 1), "hbm_frac": round(gbs32 / HBM_PEAK_GBS, 4),
        "bytes_per_iter": int(bytes_it), "kernels": k32,
        "fp64": fp64,
        "speedup_mixed_vs_fp64": round(fp64["time_s"] / (ms32 / 1e3), 3) if fp64 else None,
        "err_vs_fp64": (float(np.linalg.norm(x32 - x64) / np.linalg.norm(x64))
                        if fp64 else None),
        "roofline": roofline, "cpu_baseline": cpu, "reference_parity": parity,
    }
    print(json.dumps(line), flush=True)


def check_solves(label, results, tol, true_resid):
    """Every timed solve converged, met tol (on the FP64 true residual when
    true_resid) and took the same number of iterations; otherwise no line."""
    for r in results:
        if not r.converged or r.breakdown:
            refuse(f"{label} solve did not converge ({r.iters} iterations, rel {r.rel_resid:.3e})")
        if not (r.rel_resid < tol) or (true_resid and not (r.true_rel_resid < tol)):
            refuse(f"{label} solve missed tol {tol:g}: rel {r.rel_resid:.3e}, true "
                   f"{r.true_rel_resid:.3e}")
    if len({r.iters for r in results}) != 1:
        refuse(f"{label} solves disagree: iterations {[r.iters for r in results]}")


def reference_parity(cpu, x_ref, r64, x64, x32):
    """The GPU solves against the reference's CG run to tolerance on the
    same system: iterations (FP64 within one, or no line), final residual
    and relative errors of x."""
    if not cpu["converged"]:
        refuse("the reference's CG did not converge")
    nr = float(np.linalg.norm(x_ref))
    out = {"reference_iters": cpu["iters"], "reference_rel_resid": cpu["rel_resid"],
           "err_mixed_vs_reference": float(np.linalg.norm(x32 - x_ref) / nr)}
    if r64 is not None:
        d = r64.iters - cpu["iters"]
        out.update({"fp64_iters": r64.iters, "fp64_rel_resid": r64.rel_resid,
                    "fp64_iters_minus_reference": d,
                    "fp64_rel_resid_minus_reference": r64.rel_resid - cpu["rel_resid"],
                    "err_fp64_vs_reference": float(np.linalg.norm(x64 - x_ref) / nr)})
        log(f"reference parity: iterations {r64.iters} vs {cpu['iters']}, rel "
            f"{r64.rel_resid:.9e} vs {cpu['rel_resid']:.9e}, ||x64-x_ref||/||x_ref|| "
            f"{out['err_fp64_vs_reference']:.3e}, mixed {out['err_mixed_vs_reference']:.3e}")
        if abs(d) > 1:
            refuse(f"FP64 iterations {r64.iters} differ from the reference's {cpu['iters']} by "
                   "more than one")
    return out


if __name__ == "__main__":
    main()
