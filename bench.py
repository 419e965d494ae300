"""Time-to-solution benchmark on one MI355X.

Workload (default): 3D jump-coefficient diffusion, 7-point finite-volume stencil
on a 160^3 grid (n = 4,096,000, nnz = 28,518,400), kappa = 10^u with u uniform on
{-3, 0, 3} per 16^3 block, harmonic-mean faces, b ~ N(0,1).  Measured path:
FP64 FGMRES(40) with two-level RAS (256 subdomains, ufactor 300, overlap 1,
FP32 ILU(0) factors and sweeps, Nicolaides coarse space with a 256 x 256 FP64
LU), CGS2, stopping on an FP64-verified ||b - A x|| / ||b|| <= 1e-8.  Beside it,
on the same GPU and matrix: plain FP64 GMRES(40) with one-level RAS and MGS
(the reference's method).  A "step" is one solve from x0 = 0 with the inputs
resident in HBM; ``value`` is seconds per solve (lower is better).

The dominant kernel family's achieved bandwidth is measured live with HIP
events on the solver stream; the CPU oracle (the reference algorithm restated
in C++) is timed on a bounded number of iterations on the host cores.

    python bench.py [--steps K] [--warmup W] [--config c3|c2|c1|c4|c5] [--no-baseline] [--no-cpu]
                    [--graph 0|1] [--no-learned]
"""
import argparse
import ctypes as C
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
import __graft_entry__  # noqa: E402

HBM_PEAK_GBS = 8000.0


def load_problems():
    spec = importlib.util.spec_from_file_location(
        "sgamd_problems", os.path.join(__graft_entry__.PKG_DIR, "problems.py"))
    P = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(P)
    return P


CONFIGS = {
    "c3": dict(workload="diffusion3d_jump_160^3", restart=40, nparts=256, overlap=1,
               ufactor=300, weighted=True, max_iter=10000, base_max_iter=12000),
    "c2": dict(workload="poisson3d_128^3", restart=50, nparts=64, overlap=2, ufactor=0,
               weighted=False, max_iter=2000, base_max_iter=4000),
    "c1": dict(workload="poisson2d_256^2", restart=30, nparts=1, overlap=0, ufactor=0,
               weighted=False, max_iter=2000, base_max_iter=2000),
    "c5": dict(workload="diffusion2d_grf_1024^2", restart=30, nparts=256, overlap=2, ufactor=0,
               weighted=True, max_iter=10000, base_max_iter=20000),
    "c4": dict(workload="batch_64x_diffusion2d_grf", restart=30, nparts=0, overlap=2, ufactor=0,
               weighted=False, max_iter=2000, base_max_iter=4000),
}


def make_problem(P, cfg):
    if cfg == "c3":
        k = P.block_jump_kappa(160, 16, 3)
        rp, ci, v = P.diffusion3d(k)
        b = P.normal(33, len(rp) - 1)
        xs = None
    elif cfg == "c5":
        rp, ci, v = P.diffusion2d(np.exp(P.grf(1024, 1.5, 0.05, 5)))
        b = P.normal(55, len(rp) - 1)
        xs = None
    elif cfg == "c2":
        rp, ci, v = P.poisson3d(128, 128, 128)
        xs = P.uniform_pm1(2, len(rp) - 1)
        b = None
    else:
        rp, ci, v = P.poisson2d(256, 256)
        xs = P.uniform_pm1(1, len(rp) - 1)
        b = None
    return rp, ci, v, b, xs


def spmv_np(rp, ci, v, x):
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    return np.bincount(rows, weights=v * x[ci], minlength=len(rp) - 1)


def run_solves(sg, ctx, A, M, db, dx, k, **kw):
    out = []
    for _ in range(k):
        dx.zero()
        ctx.sync()
        t0 = time.perf_counter()
        r = sg.solve(A, db, dx, M, **kw)
        ctx.sync()
        r["wall_s"] = time.perf_counter() - t0
        out.append(r)
    return out


KERNEL_NAMES = {"spmv": "k_spmv_*", "trsv": "k_sweep_*", "gram_dot": "k_gram_dot",
                "gram_update": "k_gram_update"}


def traffic_models(n, nnz, info, restart, fp32):
    """Algorithmic bytes per launch of the SpMV and the fused RAS sweep."""
    v = 4 if fp32 else 8
    return {
        # SpMV (FP64 CSR): vals+cols, row_ptr, x, y
        "spmv": 12.0 * nnz + 20.0 * n + 4,
        # fused RAS apply: factors+cols, level/row metadata, gather index + x, scatter + y
        "trsv": (v + 4.0) * info["nnz_ext"] + 8.0 * info["n_ext"] + 12.0 * info["n_ext"]
                + 12.0 * n,
    }


def measured_traffic(workload, kernel):
    """HBM bytes per launch of a kernel from the committed rocprofv3 counter
    summary (FETCH_SIZE / WRITE_SIZE, calibrated on a plain copy); None when
    this workload has not been profiled."""
    f = os.path.join(ROOT, "profiles", "traffic_c3_diffusion160.json")
    try:
        with open(f) as fh:
            t = json.load(fh)
    except OSError:
        return None
    if t.get("workload") != workload or kernel not in t.get("kernels", {}):
        return None
    return float(t["kernels"][kernel]["bytes"])


def kernel_rooflines(ctx, workload, n, nnz, info, fp32):
    """Roofline entry of each hot kernel of the last profiled solve: HIP-event
    time per launch on the solver stream against its algorithmic bytes per
    launch (Gram passes: 8 n (j + 2) and 8 n (j + 3) for the j of each launch,
    from the columns the library recorded)."""
    models = traffic_models(n, nnz, info, 0, fp32)
    out = {}
    for f in ("spmv", "trsv", "gram_dot", "gram_update"):
        launches, ms = ctx.kernel_stats(f)
        if not launches:
            continue
        if f == "gram_dot":
            total = 8.0 * n * (ctx.kernel_units(f) + launches)
        elif f == "gram_update":
            total = 8.0 * n * (ctx.kernel_units(f) + 2 * launches)
        else:
            total = models[f] * launches
        avg_s = ms * 1e-3 / launches
        bpl = total / launches
        ach = bpl / avg_s / 1e9
        out[f] = {"bound": "hbm", "kernel": KERNEL_NAMES[f], "achieved": round(ach, 1),
                  "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                  "traffic": measured_traffic(workload, f), "bytes_per_launch": round(bpl),
                  "avg_launch_us": round(avg_s * 1e6, 2), "launches": launches,
                  "total_ms": round(ms, 3)}
    return out


def cpu_baseline(O, rp, ci, v, b, part, cfg, iters, est_iters, est_label):
    """The oracle's FP64 FGMRES + two-level RAS (the reference algorithm, OpenMP
    over subdomains as in src/precond/ras.h:48) on the host cores: setup timed
    once, then the median over 3 runs of the time per iteration of `iters`
    iterations; value = setup + that rate x est_iters (an extrapolation to the
    iteration count of the same FP64 method, est_label)."""
    A = O.Csr(rp, ci, v)
    t0 = time.perf_counter()
    pc = O.Precond("twolevel_ras", A, cfg["nparts"], cfg["overlap"], part, precision=0)
    t_setup = time.perf_counter() - t0
    rates = []
    for _ in range(3):
        t0 = time.perf_counter()
        r = O.solve("fgmres", A, b, pc, cfg["restart"], iters, 0.0)
        rates.append((time.perf_counter() - t0) / max(1, r["iterations"]))
    t_iter = float(np.median(rates))
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(t_setup + t_iter * est_iters, 3), "unit": "s",
            "cores": threads, "kind": "port",
            "setup_s": round(t_setup, 3), "s_per_iteration": round(t_iter, 5),
            "sample": f"measured: setup {t_setup:.2f} s and the median of 3 runs of "
                      f"{r['iterations']} FGMRES({cfg['restart']}) iterations of the oracle's "
                      f"FP64 two-level RAS on the same matrix and partition "
                      f"({t_iter * 1e3:.1f} ms/iteration); extrapolated: value = setup + "
                      f"{est_iters} iterations ({est_label}) at that rate"}


def cpu_full(args, sg, P, cfg):
    """The oracle's FP64 two-level RAS-FGMRES (the reference algorithm, OpenMP
    over subdomains) run to ||r|| <= 1e-8 ||b|| on the host cores, with no GPU:
    the measured CPU time-to-solution beside the bench's per-iteration
    extrapolation.  Runs in pieces of ten restart cycles, each continuing from
    the last x (a restart happens there anyway, so the iterates are those of
    one run), printing a progress line per piece."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as O
    rp, ci, v, b, xs = make_problem(P, args.config)
    if b is None:
        b = spmv_np(rp, ci, v, xs)
    t0 = time.perf_counter()
    part = sg.partition(rp, ci, cfg["nparts"], cfg["ufactor"], vals=v, weighted=cfg["weighted"])
    t_part = time.perf_counter() - t0
    A = O.Csr(rp, ci, v)
    t0 = time.perf_counter()
    pc = O.Precond("twolevel_ras", A, cfg["nparts"], cfg["overlap"], part, precision=0)
    t_setup = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    print(f"# cpu full: partition {t_part:.1f} s, oracle setup {t_setup:.1f} s, "
          f"{threads} threads", file=sys.stderr, flush=True)
    tol = 1e-8 * float(np.linalg.norm(b))
    x = np.zeros(len(b))
    its, t_solve, conv = 0, 0.0, False
    piece = 10 * cfg["restart"]
    while its < cfg["base_max_iter"] and not conv:
        t0 = time.perf_counter()
        r = O.solve("fgmres", A, b, pc, cfg["restart"], piece, tol, x0=x)
        t_solve += time.perf_counter() - t0
        x, its, conv = r["x"], its + r["iterations"], r["converged"]
        print(f"# cpu full: {its} it, {t_solve:.0f} s, |r| {r['residual_norm']:.3e} "
              f"(target {tol:.3e})", file=sys.stderr, flush=True)
    rel = float(np.linalg.norm(b - spmv_np(rp, ci, v, x)) / np.linalg.norm(b))
    print(json.dumps({"cpu_full": {
        "workload": cfg["workload"], "kind": "port", "cores": threads,
        "solver": f"fgmres({cfg['restart']})+twolevel_ras fp64",
        "setup_s": round(t_setup, 3), "partition_s": round(t_part, 3),
        "solve_s": round(t_solve, 3), "value": round(t_setup + t_solve, 3), "unit": "s",
        "iterations": its, "converged": conv, "rel_true_residual": rel,
        "s_per_iteration": round(t_solve / max(1, its), 5)}}))


def probe(sg, ctx, A, M, db, dx, k, info, n, nnz):
    """K preconditioner applies and K SpMVs with per-launch HIP-event timing."""
    dy = sg.DeviceArray(ctx, n)
    M.apply(db, dy)  # first apply runs the shape autotuner
    A.spmv(db, dy)
    ctx.sync()
    ctx.profiling(True)
    ctx.reset_stats()
    for _ in range(k):
        M.apply(db, dy)
        A.spmv(dy, dx)
    ctx.sync()
    fams = {f: ctx.kernel_stats(f) for f in ("spmv", "trsv", "precond")}
    ctx.profiling(False)
    models = traffic_models(n, nnz, info, 0, True)
    out = {f: {"launches": fams[f][0], "avg_us": round(1e3 * fams[f][1] / max(1, fams[f][0]), 2)}
           for f in fams}
    for f in ("spmv", "trsv"):
        out[f]["GB_s"] = round(models[f] / (out[f]["avg_us"] * 1e-6) / 1e9, 1)
    out["precond_info"] = info
    print(json.dumps(out))


def sweep_clock(sg, ctx, M, db, n, k, out_path=None):
    """K preconditioner applies with the staged sweep's workgroup clock armed.

    Per apply: span = last workgroup end - first workgroup start; busy = the
    sum of workgroup spans / (256 CUs x span), so 1 - busy is the CU time the
    sweep leaves idle; balanced = sum of spans / 256, the span if the same
    work were spread evenly.  A least-squares fit span_s ~ a rows_s + b
    levels_s + c over the subdomains says which size the span follows."""
    dy = sg.DeviceArray(ctx, n)
    M.apply(db, dy)  # first apply runs the shape autotuner
    ctx.sync()
    recs = [M.sweep_clock(db, dy) for _ in range(k)]
    cus = 256
    rows = []
    for r in recs:
        d = r["end_us"] - r["start_us"]
        span = float(r["end_us"].max())
        rows.append((span, float(d.sum() / (cus * span)), float(d.sum() / cus), float(d.max()),
                     float(np.median(d)), float(r["start_us"].max())))
    a = np.array(rows)
    med = np.median(a, axis=0)
    r = recs[int(np.argsort(a[:, 0])[len(a) // 2])]  # the median apply
    d = r["end_us"] - r["start_us"]
    lev = (r["flev"] + r["blev"]).astype(float)
    X = np.stack([r["rows"].astype(float), lev, np.ones_like(lev)], 1)
    coef, *_ = np.linalg.lstsq(X, d, rcond=None)
    fit = X @ coef
    r2 = 1 - float(np.sum((d - fit) ** 2) / np.sum((d - d.mean()) ** 2))
    per_cu = np.unique(r["hwid"], return_counts=True)[1]
    t = int(np.argmax(d))
    ph = {"gather": r["gathered_us"] - r["start_us"], "sweeps": r["swept_us"] - r["gathered_us"],
          "scatter": r["end_us"] - r["swept_us"]}
    out = {"sweep_clock": {
        "applies": k, "workgroups": int(len(d)), "span_us": round(med[0], 2),
        "cu_busy_frac": round(med[1], 4), "balanced_us": round(med[2], 2),
        "longest_us": round(med[3], 2), "median_wg_us": round(med[4], 2),
        "last_start_us": round(med[5], 2),
        "distinct_cus": int(len(per_cu)), "max_wg_per_cu": int(per_cu.max()),
        "longest": {"sub": int(r["sub"][t]), "rows": int(r["rows"][t]),
                    "levels": int(lev[t]), "chunks": int(r["chunks"][t]),
                    "phases_us": {k2: round(float(v2[t]), 2) for k2, v2 in ph.items()}},
        "median_phases_us": {k2: round(float(np.median(v2)), 2) for k2, v2 in ph.items()},
        "max_phases_us": {k2: round(float(v2.max()), 2) for k2, v2 in ph.items()},
        "fit_us": {"per_krow": round(1e3 * coef[0], 3), "per_level": round(coef[1], 4),
                   "const": round(coef[2], 2), "r2": round(r2, 3)},
        "rate_khz": r["rate_khz"]}}
    print(json.dumps(out))
    if out_path:
        np.savez(out_path, **{k2: v2 for k2, v2 in r.items() if k2 != "rate_khz"})


def copy_probe():
    import torch
    a = torch.ones(32 << 20, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": a.numel() * 8, "copy_probe_bytes_written":
                      b.numel() * 8, "launches": 5}))


def run_batch(args, sg, P, cfg):
    """64 independent 2D GRF-diffusion systems solved together (one graph per
    restart cycle) vs each system solved alone with FP64 RAS-GMRES."""
    t0 = time.perf_counter()
    systems = P.batch_systems(64, base_seed=4000)
    mats = [(rp, ci, v) for rp, ci, v, _ in systems]
    nparts = [P.batch_nparts(len(rp) - 1) for rp, _, _ in mats]
    part = sg.batch_partition(mats, nparts)
    b = np.concatenate([s[3] for s in systems])
    print(f"# batch: {len(systems)} systems, n={len(b)}, sides "
          f"{min(int(np.sqrt(len(s[3]))) for s in systems)}..."
          f"{max(int(np.sqrt(len(s[3]))) for s in systems)} ({time.perf_counter() - t0:.1f} s)",
          file=sys.stderr, flush=True)
    ctx = sg.Context(0)
    B = sg.BatchCsr(ctx, mats)
    t0 = time.perf_counter()
    M = sg.Precond(ctx, "ras", nparts=int(sum(nparts)), overlap=cfg["overlap"],
                   precision="fp32", partition=part).setup(B)
    ctx.sync()
    t_setup = time.perf_counter() - t0
    db = sg.DeviceArray.from_host(ctx, b)
    dx = sg.DeviceArray(ctx, B.n)
    kw = dict(method="fgmres", restart=cfg["restart"], max_iter=cfg["max_iter"], rtol=1e-8,
              ortho="cgs2", graph=bool(args.graph))
    times, rs = [], None
    for it in range(args.warmup + args.steps):
        dx.zero()
        ctx.sync()
        t = time.perf_counter()
        rs = sg.solve_batch(B, db, dx, M, **kw)
        ctx.sync()
        if it >= args.warmup:
            times.append(time.perf_counter() - t)
    x = dx.get()
    rels = []
    for q, (rp, ci, v, bq) in enumerate(systems):
        xq = x[B.sys_rows[q]:B.sys_rows[q + 1]]
        rels.append(np.linalg.norm(bq - spmv_np(rp, ci, v, xq)) / np.linalg.norm(bq))
    assert all(r["converged"] for r in rs) and max(rels) <= 1e-8 * 1.0001, max(rels)
    its = [r["iterations"] for r in rs]
    print(f"# batched fgmres+ras fp32: {np.mean(times):.4f} s, iterations "
          f"{min(its)}..{max(its)}, max rel {max(rels):.2e}", file=sys.stderr, flush=True)
    # solution error against an FP64 reference: the same batch with the FP64
    # preconditioner to ||b - A x|| / ||b|| <= 1e-12 per system (skipped with
    # --no-baseline, as the other reference legs are)
    xerr, extra = [], {}
    if not args.no_baseline:
        M64 = sg.Precond(ctx, "ras", nparts=int(sum(nparts)), overlap=cfg["overlap"],
                         precision="fp64", partition=part).setup(B)
        dx.zero()
        rr = sg.solve_batch(B, db, dx, M64, **dict(kw, rtol=1e-12, max_iter=4 * cfg["max_iter"]))
        ctx.sync()
        xref = dx.get()
        del M64
        rref = []
        for q, (rp, ci, v, bq) in enumerate(systems):
            lo, hi = B.sys_rows[q], B.sys_rows[q + 1]
            xerr.append(float(np.linalg.norm(x[lo:hi] - xref[lo:hi]) / np.linalg.norm(xref[lo:hi])))
            rref.append(np.linalg.norm(bq - spmv_np(rp, ci, v, xref[lo:hi])) / np.linalg.norm(bq))
        extra["x_reference"] = {"converged": all(r["converged"] for r in rr),
                                "iterations": f"{min(r['iterations'] for r in rr)}.."
                                              f"{max(r['iterations'] for r in rr)}",
                                "max_rel_true_residual": float(max(rref)),
                                "what": "the same batch with the FP64 preconditioner, to 1e-12"}
    if not args.no_baseline:
        tb, bits = 0.0, []
        for (rp, ci, v, bq), pq, k in zip(systems, np.split(part, B.sys_rows[1:-1]), nparts):
            A = sg.Csr(ctx, rp, ci, v)
            Mb = sg.Precond(ctx, "ras", nparts=k, overlap=cfg["overlap"], precision="fp64",
                            partition=pq - pq.min()).setup(A)
            dbq = sg.DeviceArray.from_host(ctx, bq)
            dxq = sg.DeviceArray(ctx, len(bq))
            r = run_solves(sg, ctx, A, Mb, dbq, dxq, 1, method="gmres", restart=cfg["restart"],
                           max_iter=cfg["base_max_iter"], rtol=1e-8, ortho="mgs")[0]
            tb += r["wall_s"]
            bits.append(r["iterations"])
            del Mb, A
        extra["fp64_ras_gmres_sequential"] = {"s": round(tb, 4), "iterations_min": min(bits),
                                              "iterations_max": max(bits)}
        extra["speedup_vs_fp64_ras_gmres"] = round(tb / np.mean(times), 2)
        print(f"# baseline: 64 sequential FP64 RAS-GMRES solves {tb:.3f} s", file=sys.stderr,
              flush=True)
    out = {
        "metric": "time to ||r||/||b||<=1e-8 (s): mixed/learned FGMRES vs FP64 RAS-GMRES, 1 GPU",
        "value": round(float(np.mean(times)), 5), "unit": "s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * float(np.mean(times)), 3), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64 outer / f32 preconditioner",
        "data": "synthetic",
        "config": {"workload": cfg["workload"], "systems": len(systems), "n": int(B.n),
                   "nnz": B.nnz, "solver": f"batched fgmres({cfg['restart']})+ras",
                   "nparts": f"{min(nparts)}..{max(nparts)} per system", "overlap": cfg["overlap"],
                   "iterations": f"{min(its)}..{max(its)}", "max_rel_true_residual": max(rels),
                   "max_x_rel_error": max(xerr) if xerr else None,
                   "median_x_rel_error": float(np.median(xerr)) if xerr else None,
                   "x_rel_error_vs": "FP64-preconditioned batch solution at 1e-12",
                   "setup_s": round(t_setup, 3), "graph": bool(args.graph),
                   "parallelism": "single GPU"},
        "roofline": None, "cpu_baseline": None,
    }
    out.update(extra)
    print(json.dumps(out))


SHIPPED_WEIGHTS = "unet_grf1024_c4.npy"


def learned_leg(args, sg, P, cfg, with_baseline):
    """2D GRF diffusion 1024^2: mixed two-level RAS-FGMRES(30) with the learned
    stage (network on 4 x 4 block averages by default) against the same solver
    without it (the ablation), optionally also FP64 RAS-GMRES; returns a dict."""
    t0 = time.perf_counter()
    side = 1024
    g = P.grf(side, 1.5, 0.05, 5)
    rp, ci, v = P.diffusion2d(np.exp(g))
    b = P.normal(55, side * side)
    n, nnz = len(rp) - 1, int(rp[-1])
    print(f"# problem diffusion2d_grf_1024^2: n={n} nnz={nnz} ({time.perf_counter() - t0:.1f} s)",
          file=sys.stderr, flush=True)
    ctx = sg.Context(0)
    ctx.autotune(args.autotune)
    A = sg.Csr(ctx, rp, ci, v)
    db = sg.DeviceArray.from_host(ctx, b)
    dx = sg.DeviceArray(ctx, n)
    t0 = time.perf_counter()
    part = sg.partition(rp, ci, cfg["nparts"], cfg["ufactor"], vals=v, weighted=cfg["weighted"])
    base = sg.Precond(ctx, "twolevel_ras", nparts=cfg["nparts"], overlap=cfg["overlap"],
                      precision="fp32", partition=part, coarse_solve=1).setup(A)
    ctx.sync()
    t_setup = time.perf_counter() - t0
    if getattr(args, "sweep_clock", 0):
        sweep_clock(sg, ctx, base, db, n, args.sweep_clock, args.clock_out)
        return None
    kw = dict(method="fgmres", restart=cfg["restart"], max_iter=cfg["max_iter"], rtol=1e-8,
              ortho="cgs2")

    def timed(M, label):
        run_solves(sg, ctx, A, M, db, dx, max(1, args.warmup), **kw)
        rs = run_solves(sg, ctx, A, M, db, dx, max(3, args.steps), **kw)
        x = dx.get()
        rel = np.linalg.norm(b - spmv_np(rp, ci, v, x)) / np.linalg.norm(b)
        r = rs[-1]
        assert r["converged"] and rel <= 1e-8 * 1.0001, (label, r, rel)
        t = float(np.median([q["wall_s"] for q in rs]))
        print(f"# {label}: {r['iterations']} it, {t:.4f} s, rel {rel:.2e}", file=sys.stderr,
              flush=True)
        return t, r["iterations"], rel, x

    t_plain, it_plain, rel_p, x_plain = timed(base, "plain mixed two-level RAS-FGMRES")
    out = {"workload": "diffusion2d_grf_1024^2", "n": n, "nnz": nnz,
           "setup_s": round(t_setup, 3),
           "plain_mixed_twolevel_fgmres": {"s": round(t_plain, 4), "iterations": it_plain,
                                           "rel_true_residual": rel_p}}
    wpath = args.weights or os.path.join(__graft_entry__.PKG_DIR, "weights", SHIPPED_WEIGHTS)
    net = sg.UNet(ctx, np.load(wpath))
    stage = sg.LearnedStage(ctx, base, net, g.ravel(), side, side, min_gain=args.min_gain,
                            every=args.every, coarsen=args.coarsen).setup(A)
    t_l, it_l, rel_l, x_l = timed(stage, "learned stage FGMRES")
    calls, acc = stage.counts()
    out["learned_stage"] = {"s": round(t_l, 4), "iterations": it_l, "rel_true_residual": rel_l,
                            "coarsen": args.coarsen, "every": args.every,
                            "stage_calls": calls, "stage_accepted": acc,
                            "weights": os.path.basename(wpath)}
    out["learned_speedup_vs_plain"] = round(t_plain / t_l, 3)
    # solution error against an FP64 reference: FP64 two-level RAS-FGMRES on the
    # same partition to ||b - A x|| / ||b|| <= 1e-12 (skipped with --no-baseline)
    if not getattr(args, "no_baseline", False):
        M64 = sg.Precond(ctx, "twolevel_ras", nparts=cfg["nparts"], overlap=cfg["overlap"],
                         precision="fp64", partition=part, coarse_solve=1).setup(A)
        rr = run_solves(sg, ctx, A, M64, db, dx, 1, **dict(kw, rtol=1e-12,
                                                           max_iter=4 * cfg["max_iter"]))[0]
        xref = dx.get()
        nref = np.linalg.norm(xref)
        out["x_reference"] = {
            "iterations": rr["iterations"], "converged": rr["converged"],
            "rel_true_residual": float(np.linalg.norm(b - spmv_np(rp, ci, v, xref))
                                       / np.linalg.norm(b)),
            "what": "FP64 two-level RAS-FGMRES, same partition, to 1e-12"}
        out["plain_mixed_twolevel_fgmres"]["x_rel_error"] = float(
            np.linalg.norm(x_plain - xref) / nref)
        out["learned_stage"]["x_rel_error"] = float(np.linalg.norm(x_l - xref) / nref)
        del M64
    if with_baseline:
        Mb = sg.Precond(ctx, "ras", nparts=cfg["nparts"], overlap=cfg["overlap"],
                        precision="fp64", partition=part).setup(A)
        rb = run_solves(sg, ctx, A, Mb, db, dx, 1, method="gmres", restart=cfg["restart"],
                        max_iter=cfg["base_max_iter"], rtol=1e-8, ortho="mgs")[0]
        out["fp64_ras_gmres"] = {"s": round(rb["wall_s"], 4), "iterations": rb["iterations"],
                                 "converged": rb["converged"]}
        out["speedup_vs_fp64_ras_gmres"] = round(rb["wall_s"] / t_l, 2)
        print(f"# baseline gmres+ras fp64: {rb['iterations']} it {rb['wall_s']:.3f} s",
              file=sys.stderr, flush=True)
    del stage, net, base
    return out


def run_learned(args, sg, P, cfg):
    """The 1024^2 learned-stage workload as the bench line."""
    leg = learned_leg(args, sg, P, cfg, not args.no_baseline)
    if leg is None:  # --sweep-clock: the base preconditioner's clock only
        return
    value = leg["learned_stage"]["s"]
    out = {
        "metric": "time to ||r||/||b||<=1e-8 (s): mixed/learned FGMRES vs FP64 RAS-GMRES, 1 GPU",
        "value": round(value, 5), "unit": "s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * value, 3), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 outer / f32 preconditioner / bf16 network", "data": "synthetic",
        "config": {"workload": cfg["workload"], "n": leg["n"], "nnz": leg["nnz"],
                   "solver": f"fgmres({cfg['restart']})+twolevel_ras+learned",
                   "nparts": cfg["nparts"], "overlap": cfg["overlap"],
                   "setup_s": leg["setup_s"], "parallelism": "single GPU"},
        "roofline": None, "cpu_baseline": None,
    }
    out.update({k: v for k, v in leg.items()
                if k not in ("workload", "n", "nnz", "setup_s")})
    print(json.dumps(out))


def unet_probe(args, sg, P, k):
    """The learned stage's network with the shipped weights: K forward passes
    at 256^2 (the grid it runs on in the 1024^2 workload, coarsen 4) and at
    1024^2 on that workload's input (normalised random residual, log kappa),
    and K/4 training steps (forward, residual-loss backward, Adam) on batches
    of 16 at 256^2; mean time of each and the forward's fraction of the bf16
    dense MFMA peak."""
    ctx = sg.Context(0)
    wpath = args.weights or os.path.join(__graft_entry__.PKG_DIR, "weights", SHIPPED_WEIGHTS)
    net = sg.UNet(ctx, np.load(wpath))
    out = {}
    for side in (256, 1024):
        g = P.grf(side, 1.5, 0.05, 5)
        r = P.normal(77, side * side)
        x = np.stack([r / np.sqrt(np.mean(r * r)), g.ravel()]).astype(np.float32)
        dx = sg.DeviceArray.from_host(ctx, x.ravel())
        dy = sg.DeviceArray(ctx, side * side, np.float32)
        for _ in range(3):
            net.forward_dev(1, side, side, dx, dy)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(k):
            net.forward_dev(1, side, side, dx, dy)
        ctx.sync()
        t = (time.perf_counter() - t0) / k
        flops = unet_flops(side, side)
        out[f"forward_{side}"] = {"ms": round(t * 1e3, 4), "gflop": round(flops / 1e9, 2),
                                  "tflop_s": round(flops / t / 1e12, 1),
                                  "frac_bf16_peak": round(flops / t / 2.5e15, 4)}
    # training steps: B = 16 samples of 256^2 (residual loss, as learned.py trains)
    B, side = 16, 256
    rng = np.random.default_rng(3)
    xb = np.stack([np.stack([rng.standard_normal((side, side)),
                             P.grf(side, 1.5, 0.05, 100 + q)]) for q in range(B)]).astype(np.float32)
    dxb = sg.DeviceArray.from_host(ctx, xb.ravel())
    net.set_loss(1)
    kt = max(1, k // 4)
    for _ in range(2):
        net.loss_grad_dev(B, side, side, dxb, None)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(kt):
        net.loss_grad_dev(B, side, side, dxb, None)
        net.adam(1e-6)
    ctx.sync()
    t = (time.perf_counter() - t0) / kt
    out["train_step_16x256"] = {"ms": round(t * 1e3, 3), "steps": kt}
    out["forwards"] = k
    print(json.dumps(out))


def unet_flops(H, W):
    """Multiply-add flops of one forward: 2 * 9 * cin * cout per pixel of each
    3x3 convolution at its level (the 1x1 head and pooling are left out)."""
    spec = [(2, 16, 0), (16, 16, 0), (16, 32, 1), (32, 32, 1), (32, 64, 2), (64, 64, 2),
            (64, 64, 3), (64, 64, 3), (128, 64, 2), (64, 64, 2), (96, 32, 1), (32, 32, 1),
            (48, 16, 0), (16, 16, 0)]
    return sum(2.0 * 9 * ci * co * (H >> lv) * (W >> lv) for ci, co, lv in spec)


def _timed(fn, k):
    out, ts = None, []
    for _ in range(k):
        t0 = time.perf_counter()
        out = fn()
        ts.append(time.perf_counter() - t0)
    return out, float(np.mean(ts))


def _families(sg, ctx, fn):
    ctx.profiling(True)
    ctx.reset_stats()
    fn()
    fams = {f: ctx.kernel_stats(f) for f in ("spmv", "trsv", "precond", "ortho", "update")}
    ctx.profiling(False)
    return {f: [int(a), round(float(b), 3)] for f, (a, b) in fams.items()}


def run_rows(args, sg, P):
    """The rows around the headline path -- BiCGSTAB, ILUT, the convergence
    analytics and the Matrix Market reader -- each on the GPU with the CPU
    restatement of the reference (or the reference's own reader) timed beside
    it.  One JSON line per row; times are seconds per call with inputs in HBM."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as O
    ctx = sg.Context(0)
    steps = max(1, args.steps)
    groups = set(args.row or ["c1", "c3", "mm"])

    def emit(d):
        print(json.dumps(d), flush=True)

    if "c1" in groups:
        rows_c1(args, sg, P, O, ctx, steps, emit)
    if "c3" in groups or "mm" in groups:
        rows_c3(args, sg, P, O, ctx, steps, emit, groups)


def rows_c1(args, sg, P, O, ctx, steps, emit):
    # ---- c1 matrix: Poisson 256^2, b = A x*, x* ~ U[-1, 1] ----
    rp, ci, v = P.poisson2d(256, 256)
    n = len(rp) - 1
    b = spmv_np(rp, ci, v, P.uniform_pm1(1, n))
    A = sg.Csr(ctx, rp, ci, v)
    Ao = O.Csr(rp, ci, v)
    db, dx = sg.DeviceArray.from_host(ctx, b), sg.DeviceArray(ctx, n)
    M = sg.Precond(ctx, "ilu0").setup(A)
    Mo = O.Precond("ilu0", Ao)

    def gsolve(MM, **kw):
        def f():
            dx.zero()
            ctx.sync()
            r = sg.solve(A, db, dx, MM, **kw)
            ctx.sync()
            return r
        return f

    # BiCGSTAB + ILU(0)
    kw = dict(method="bicgstab", max_iter=2000, tol=1e-10)
    gsolve(M, **kw)()
    r, t = _timed(gsolve(M, **kw), steps)
    fams = _families(sg, ctx, gsolve(M, **kw))
    ro, tc = _timed(lambda: O.bicgstab(Ao, b, Mo, 2000, 1e-10), 1)
    vec_ms = fams["update"][1] + fams["ortho"][1]
    vec_bytes = 152.0 * n * r["iterations"]  # 19 vector passes of 8 bytes per iteration
    emit({"row": "bicgstab", "workload": "poisson2d_256^2", "precond": "ilu0", "tol": 1e-10,
          "gpu_s": round(t, 5), "iterations": r["iterations"],
          "rel_true_residual": r["rel_true_residual"],
          "cpu_s": round(tc, 4), "cpu_iterations": ro["iterations"], "cpu_kind": "port",
          "cpu_cores": 1, "speedup": round(tc / t, 2), "kernel_ms": fams,
          "vector_kernels_GB_s": round(vec_bytes / (vec_ms * 1e-3) / 1e9, 1) if vec_ms else None})

    # GMRES(30) + ILUT(1e-4, 10)
    t0 = time.perf_counter()
    Mt = sg.Precond(ctx, "ilut", drop_tol=1e-4, fill_factor=10).setup(A)
    ctx.sync()
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    Mto = O.Precond("ilut", Ao, drop_tol=1e-4, fill_factor=10)
    t_setup_cpu = time.perf_counter() - t0
    kw = dict(method="gmres", restart=30, max_iter=2000, tol=1e-10, ortho="mgs")
    gsolve(Mt, **kw)()
    r, t = _timed(gsolve(Mt, **kw), steps)
    fams = _families(sg, ctx, gsolve(Mt, **kw))
    ro, tc = _timed(lambda: O.solve("gmres", Ao, b, Mto, 30, 2000, 1e-10), 1)
    info = Mt.stats()
    emit({"row": "ilut", "workload": "poisson2d_256^2", "solver": "gmres(30)", "tol": 1e-10,
          "l_nnz_u_nnz": list(Mt.ilut_nnz()), "levels": [info["levels_fwd"], info["levels_bwd"]],
          "setup_s": round(t_setup, 3), "cpu_setup_s": round(t_setup_cpu, 3),
          "gpu_s": round(t, 5), "iterations": r["iterations"],
          "rel_true_residual": r["rel_true_residual"], "cpu_s": round(tc, 4),
          "cpu_iterations": ro["iterations"], "cpu_kind": "port", "cpu_cores": 1,
          "speedup": round(tc / t, 2), "kernel_ms": fams})

    # condition estimate (80 Lanczos steps) and spectral radius (300 power steps)
    def gan():
        c = sg.estimate_condition(A, M, 80)
        s = sg.estimate_spectral_radius(A, M, 300)
        ctx.sync()
        return c, s
    gan()
    (c, sr), t = _timed(gan, steps)
    (co, so), tc = _timed(lambda: (O.condition(Ao, Mo, 80), O.spectral(Ao, Mo, 300)), 1)
    emit({"row": "analysis", "workload": "poisson2d_256^2", "precond": "ilu0",
          "gpu_s": round(t, 5), "lanczos_steps": c["lanczos_steps"], "kappa": c["kappa"],
          "power_iters": sr["power_iters"], "rho": sr["spectral_radius"],
          "cpu_s": round(tc, 4), "cpu_kappa": co["kappa"], "cpu_rho": so["rho"],
          "cpu_power_iters": so["power_iters"], "cpu_kind": "port", "cpu_cores": 1,
          "speedup": round(tc / t, 2)})
    del M, Mt, A, db, dx


def rows_c3(args, sg, P, O, ctx, steps, emit, groups):
    def gsolve(MM, **kw):
        def f():
            dx.zero()
            ctx.sync()
            r = sg.solve(A, db, dx, MM, **kw)
            ctx.sync()
            return r
        return f

    # ---- c3 matrix: BiCGSTAB with the FP64 two-level RAS (with the FP32 one,
    # whose rounding makes the operator slightly nonlinear, BiCGSTAB stagnates
    # near 1e-3 on this 10^6-contrast problem) ----
    cfg = CONFIGS["c3"]
    rp, ci, v, b, _ = make_problem(P, "c3")
    n = len(rp) - 1
    A = sg.Csr(ctx, rp, ci, v)
    db, dx = sg.DeviceArray.from_host(ctx, b), sg.DeviceArray(ctx, n)
    if "c3" not in groups:
        return rows_mm(sg, cfg, rp, ci, v, emit)
    part = sg.partition(rp, ci, cfg["nparts"], cfg["ufactor"], vals=v, weighted=True)
    M = sg.Precond(ctx, "twolevel_ras", nparts=cfg["nparts"], overlap=cfg["overlap"],
                   precision="fp64", partition=part, coarse_solve=1).setup(A)
    kw = dict(method="bicgstab", max_iter=5000, tol=1e-8)  # relative to ||r0|| = ||b||
    gsolve(M, **kw)()
    r, t = _timed(gsolve(M, **kw), steps)
    fams = _families(sg, ctx, gsolve(M, **kw))
    cpu = None
    if not args.no_cpu:
        Ao = O.Csr(rp, ci, v)
        Mo = O.Precond("twolevel_ras", Ao, cfg["nparts"], cfg["overlap"], part, precision=0)
        k = 8
        t0 = time.perf_counter()
        ro = O.bicgstab(Ao, b, Mo, k, 0.0)
        per = (time.perf_counter() - t0) / max(1, ro["iterations"])
        cpu = {"value": round(per * r["iterations"], 2), "unit": "s", "kind": "port",
               "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
               "sample": f"{ro['iterations']} BiCGSTAB iterations of the oracle's FP64 two-level "
                         f"RAS ({per * 1e3:.0f} ms/iteration), scaled to {r['iterations']}"}
    emit({"row": "bicgstab", "workload": cfg["workload"], "precond": "twolevel_ras fp64",
          "tol": 1e-8, "gpu_s": round(t, 4), "iterations": r["iterations"],
          "rel_true_residual": r["rel_true_residual"], "kernel_ms": fams, "cpu_baseline": cpu})
    (c, sr), t = _timed(lambda: (sg.estimate_condition(A, M, 80),
                                 sg.estimate_spectral_radius(A, M, 300)), 1)
    emit({"row": "analysis", "workload": cfg["workload"], "precond": "twolevel_ras fp64",
          "gpu_s": round(t, 4), "lanczos_steps": c["lanczos_steps"], "kappa": c["kappa"],
          "lambda_min": c["lambda_min"], "lambda_max": c["lambda_max"],
          "power_iters": sr["power_iters"], "rho": sr["spectral_radius"]})
    del M
    if "mm" in groups:
        rows_mm(sg, cfg, rp, ci, v, emit)


def rows_mm(sg, cfg, rp, ci, v, emit):
    # ---- Matrix Market: the c3 matrix as a 17-digit coordinate file ----
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "c3.mtx")
        sg.write_matrix_market(path, rp, ci, v, digits=17)
        size = os.path.getsize(path)
        out, t = _timed(lambda: sg.read_matrix_market(path), 1)
        assert np.array_equal(out[3], ci) and np.array_equal(out[4], v)
        ref = None
        harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
        if os.path.exists(harness):
            import subprocess
            q = subprocess.run([harness, "mmtime", path], capture_output=True, text=True,
                               timeout=600)
            if q.returncode == 0:
                ref = json.loads(q.stdout)
        emit({"row": "matrix_market", "workload": cfg["workload"], "file_MB": round(size / 1e6, 1),
              "read_s": round(t, 3), "MB_s": round(size / 1e6 / t, 1),
              "reference_read_s": ref["seconds"] if ref else None,
              "reference_kind": "reference (read_matrix_market, 1 core)" if ref else None,
              "speedup": round(ref["seconds"] / t, 2) if ref else None})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3", choices=sorted(CONFIGS))
    ap.add_argument("--no-baseline", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=15)
    ap.add_argument("--cpu-full", action="store_true",
                    help="only run the CPU oracle to convergence on the host cores (no GPU; "
                         "minutes to tens of minutes at c3), print one JSON line and exit")
    ap.add_argument("--autotune", type=int, default=1,
                    help="kernel-shape mode (sg_context_set_autotune): 1 tune, 0 defaults, "
                         "2..11 force a shape")
    ap.add_argument("--probe", type=int, default=0,
                    help="profiling aid: after setup run this many preconditioner applies and "
                         "SpMVs on the workload, print their mean times and exit")
    ap.add_argument("--sweep-clock", type=int, default=0,
                    help="profiling aid: after setup run this many preconditioner applies "
                         "with the staged sweep's workgroup clock, print CU busy time and exit")
    ap.add_argument("--clock-out", default=None,
                    help="with --sweep-clock: also save the median apply's records (.npz)")
    ap.add_argument("--weights", default=None,
                    help="c5: network parameters (.npy) of the learned stage")
    ap.add_argument("--min-gain", type=float, default=0.01)
    ap.add_argument("--every", type=int, default=1)
    ap.add_argument("--coarsen", type=int, default=4,
                    help="c5: the network sees c x c block averages (1, 2 or 4)")
    ap.add_argument("--no-learned", action="store_true",
                    help="c3: leave out the 1024^2 learned-stage leg")
    ap.add_argument("--rows", action="store_true",
                    help="measure the rows around the headline path (BiCGSTAB, ILUT, "
                         "analytics, Matrix Market reader), one JSON line each, and exit")
    ap.add_argument("--row", action="append", choices=("c1", "c3", "mm"),
                    help="with --rows: only these groups (c1: BiCGSTAB / ILUT / analytics on "
                         "Poisson 256^2; c3: BiCGSTAB / analytics on the 160^3 jump case; "
                         "mm: Matrix Market reader); repeatable")
    ap.add_argument("--unet-probe", type=int, default=0,
                    help="with --config c5: time this many network forwards (256^2 and 1024^2) "
                         "and a quarter as many training steps, then exit")
    ap.add_argument("--graph", type=int, default=1,
                    help="single systems: 1 (default) = each restart cycle captured and "
                         "replayed as a hipGraph, 0 = launched kernel by kernel")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 256 MiB device copy 5 times (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: the solver runs on a single GPU; a "
              "multi-GPU path (RCCL over xGMI) is designed but not built.",
              file=sys.stderr)
        sys.exit(1)

    sg = __graft_entry__.load_package()
    P = load_problems()
    if args.cpu_full:
        cpu_full(args, sg, P, CONFIGS[args.config])
        return
    if args.rows:
        run_rows(args, sg, P)
        return
    cfg = CONFIGS[args.config]
    if args.config == "c4":
        run_batch(args, sg, P, cfg)
        return
    if args.config == "c5":
        if args.unet_probe:
            unet_probe(args, sg, P, args.unet_probe)
            return
        run_learned(args, sg, P, cfg)
        return
    run_single(args, sg, P, cfg)


def run_single(args, sg, P, cfg):
    """One system (c3 by default; c2, c1): the measured solver, the legs that
    separate its gains, the reference method on the same GPU, per-kernel
    rooflines and the CPU oracle."""
    t0 = time.perf_counter()
    rp, ci, v, b, xs = make_problem(P, args.config)
    if b is None:
        b = spmv_np(rp, ci, v, xs)
    n, nnz = len(rp) - 1, int(rp[-1])
    print(f"# problem {cfg['workload']}: n={n} nnz={nnz} ({time.perf_counter() - t0:.1f} s)",
          file=sys.stderr, flush=True)

    ctx = sg.Context(0)
    ctx.autotune(args.autotune)
    A = sg.Csr(ctx, rp, ci, v)
    db = sg.DeviceArray.from_host(ctx, b)
    dx = sg.DeviceArray(ctx, n)
    graph = bool(args.graph)

    def true_rel(x):
        return float(np.linalg.norm(b - spmv_np(rp, ci, v, x)) / np.linalg.norm(b))

    t0 = time.perf_counter()
    t_part = 0.0
    if args.config == "c1":
        part = None
        M = sg.Precond(ctx, "ilu0").setup(A)
        method, kind = "gmres", "ilu0"
        solve_kw = dict(method="gmres", restart=cfg["restart"], max_iter=2000, rtol=1e-8,
                        ortho="cgs2", graph=graph)
    else:
        part = sg.partition(rp, ci, cfg["nparts"], cfg["ufactor"], vals=v,
                            weighted=cfg["weighted"])
        t_part = time.perf_counter() - t0
        kind = "twolevel_ras"
        M = sg.Precond(ctx, kind, nparts=cfg["nparts"], overlap=cfg["overlap"],
                       precision="fp32", partition=part, coarse_solve=1).setup(A)
        method = "fgmres"
        solve_kw = dict(method="fgmres", restart=cfg["restart"], max_iter=cfg["max_iter"],
                        rtol=1e-8, ortho="cgs2", graph=graph)
    ctx.sync()
    t_setup = time.perf_counter() - t0
    info = M.stats()
    print(f"# setup {t_setup:.2f} s (partition {t_part:.2f} s)  {info}", file=sys.stderr,
          flush=True)
    if args.sweep_clock:
        sweep_clock(sg, ctx, M, db, n, args.sweep_clock, args.clock_out)
        return
    if args.probe:
        probe(sg, ctx, A, M, db, dx, args.probe, info, n, nnz)
        return

    run_solves(sg, ctx, A, M, db, dx, args.warmup, **solve_kw)
    rs = run_solves(sg, ctx, A, M, db, dx, args.steps, **solve_kw)
    times = [r["wall_s"] for r in rs]
    r = rs[-1]
    x = dx.get()
    rel = true_rel(x)
    assert r["converged"] and rel <= 1e-8 * 1.0001, (r, rel)
    print(f"# {method}+{kind}: {r['iterations']} it, {np.mean(times):.4f} s/solve, "
          f"rel {rel:.2e}", file=sys.stderr, flush=True)

    # per-kernel time (HIP events on the solver stream), one profiled solve
    # launched kernel by kernel (events are not recorded inside a graph)
    ctx.profiling(True)
    ctx.reset_stats()
    dx.zero()
    sg.solve(A, db, dx, M, **dict(solve_kw, graph=False))
    fams = {f: ctx.kernel_stats(f) for f in ("spmv", "trsv", "precond", "ortho", "update")}
    roofs = kernel_rooflines(ctx, cfg["workload"], n, nnz, info, True)
    ctx.profiling(False)
    dom = max(roofs, key=lambda f: roofs[f]["total_ms"])
    print(f"# kernel families (launches, ms): {fams}", file=sys.stderr, flush=True)
    print(f"# kernels: " + ", ".join(f"{f} {q['avg_launch_us']} us {q['frac']}"
                                     for f, q in roofs.items()), file=sys.stderr, flush=True)

    extra = {}
    legs = {}
    x_err = None
    if args.config == "c2":
        x_err = float(np.linalg.norm(x - xs) / np.linalg.norm(xs))
    if not args.no_baseline and args.config != "c1":
        # the same solver with an FP64 preconditioner: isolates the mixed-precision factor
        M64 = sg.Precond(ctx, kind, nparts=cfg["nparts"], overlap=cfg["overlap"],
                         precision="fp64", partition=part, coarse_solve=1).setup(A)
        run_solves(sg, ctx, A, M64, db, dx, 1, **solve_kw)
        r64 = run_solves(sg, ctx, A, M64, db, dx, args.steps, **solve_kw)
        legs["fp64_twolevel_fgmres"] = {
            "s": round(float(np.mean([q["wall_s"] for q in r64])), 4),
            "iterations": r64[-1]["iterations"], "converged": r64[-1]["converged"],
            "rel_true_residual": r64[-1]["rel_true_residual"],
            "what": "same solver, partition and CGS2 with the FP64 preconditioner"}
        print(f"# fp64 {method}+{kind}: {r64[-1]['iterations']} it "
              f"{legs['fp64_twolevel_fgmres']['s']:.3f} s", file=sys.stderr, flush=True)
        if xs is None:
            # reference solution: FP64 two-level FGMRES to ||b - A x|| / ||b|| <= 1e-12
            ref_kw = dict(solve_kw, rtol=1e-12, max_iter=4 * cfg["max_iter"])
            rr = run_solves(sg, ctx, A, M64, db, dx, 1, **ref_kw)[0]
            xref = dx.get()
            legs["x_reference"] = {"iterations": rr["iterations"], "converged": rr["converged"],
                                   "rel_true_residual": true_rel(xref)}
            x_err = float(np.linalg.norm(x - xref) / np.linalg.norm(xref))
        del M64
        # the reference's partition semantics (unit edge weights, as its METIS call)
        if cfg["weighted"]:
            t0 = time.perf_counter()
            part_u = sg.partition(rp, ci, cfg["nparts"], cfg["ufactor"], vals=v, weighted=False)
            t_pu = time.perf_counter() - t0
            Mu = sg.Precond(ctx, kind, nparts=cfg["nparts"], overlap=cfg["overlap"],
                            precision="fp32", partition=part_u, coarse_solve=1).setup(A)
            ru = run_solves(sg, ctx, A, Mu, db, dx, 1, **solve_kw)[0]
            legs["unit_weight_partition"] = {
                "s": round(ru["wall_s"], 4), "iterations": ru["iterations"],
                "converged": ru["converged"], "rel_true_residual": ru["rel_true_residual"],
                "partition_s": round(t_pu, 3),
                "what": "the measured mixed solver on a unit-edge-weight partition"}
            print(f"# unit-weight partition: {ru['iterations']} it {ru['wall_s']:.3f} s "
                  f"conv={ru['converged']}", file=sys.stderr, flush=True)
            del Mu
        else:
            part_u = part
        # the reference's method: FP64 one-level RAS-GMRES with MGS, on the
        # reference's (unit-weight) partition
        Mb = sg.Precond(ctx, "ras", nparts=cfg["nparts"], overlap=cfg["overlap"],
                        precision="fp64", partition=part_u).setup(A)
        base_kw = dict(method="gmres", restart=cfg["restart"], max_iter=cfg["base_max_iter"],
                       rtol=1e-8, ortho="mgs")
        rb = run_solves(sg, ctx, A, Mb, db, dx, 1, **base_kw)[0]
        extra["fp64_ras_gmres"] = {"s": round(rb["wall_s"], 4), "iterations": rb["iterations"],
                                   "converged": rb["converged"],
                                   "rel_true_residual": rb["rel_true_residual"],
                                   "partition": "unit-weight"}
        # a baseline that hit its iteration cap gives a lower bound on the speedup
        extra["speedup_vs_fp64_ras_gmres"] = round(rb["wall_s"] / np.mean(times), 2)
        extra["speedup_is_lower_bound"] = not rb["converged"]
        print(f"# baseline gmres+ras fp64: {rb['iterations']} it {rb['wall_s']:.3f} s "
              f"conv={rb['converged']}", file=sys.stderr, flush=True)
        del Mb
        # the reference's other Krylov method on the same two-level preconditioner
        # in FP64 (BiCGSTAB needs a fixed linear operator; the FP32 one stagnates)
        Mq = sg.Precond(ctx, kind, nparts=cfg["nparts"], overlap=cfg["overlap"],
                        precision="fp64", partition=part, coarse_solve=1).setup(A)
        bq_kw = dict(method="bicgstab", max_iter=cfg["max_iter"], tol=1e-8)
        run_solves(sg, ctx, A, Mq, db, dx, 1, **bq_kw)
        rq = run_solves(sg, ctx, A, Mq, db, dx, args.steps, **bq_kw)
        legs["bicgstab_twolevel_fp64"] = {
            "s": round(float(np.mean([q["wall_s"] for q in rq])), 4),
            "iterations": rq[-1]["iterations"], "converged": rq[-1]["converged"],
            "rel_true_residual": rq[-1]["rel_true_residual"]}
        print(f"# bicgstab+{kind} fp64: {rq[-1]['iterations']} it "
              f"{legs['bicgstab_twolevel_fp64']['s']:.3f} s", file=sys.stderr, flush=True)
        del Mq

    cpu = None
    if not args.no_cpu and args.config != "c1":
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_bind as O
        if "fp64_twolevel_fgmres" in legs:
            est, lab = legs["fp64_twolevel_fgmres"]["iterations"], \
                "those of the FP64 two-level FGMRES leg on the GPU"
        else:
            est, lab = r["iterations"], "those of the measured GPU solve"
        cpu = cpu_baseline(O, rp, ci, v, b, part, cfg, args.cpu_iters, est, lab)

    out = {
        "metric": "time to ||r||/||b||<=1e-8 (s): mixed/learned FGMRES vs FP64 RAS-GMRES, 1 GPU",
        "value": round(float(np.mean(times)), 5),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * float(np.mean(times)), 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 outer / f32 preconditioner",
        "data": "synthetic",
        "config": {"workload": cfg["workload"], "n": n, "nnz": nnz,
                   "solver": f"{method}({cfg['restart']})+{kind}",
                   "nparts": cfg["nparts"], "overlap": cfg["overlap"],
                   "partition": "coefficient-weighted" if cfg["weighted"] else "unit-weight",
                   "ufactor": cfg["ufactor"],
                   "coarse": "FP64 A0 and LU on the device at setup; applied as the FP32 "
                             "explicit inverse (one dense mat-vec)" if kind != "ilu0" else None,
                   "ortho": "cgs2", "graph": graph,
                   "iterations": r["iterations"], "rel_true_residual": rel,
                   "x_rel_error": x_err,
                   "x_rel_error_vs": ("x*" if xs is not None else
                                      "FP64 two-level FGMRES solution at 1e-12"),
                   "setup_s": round(t_setup, 3), "partition_s": round(t_part, 3),
                   "parallelism": "single GPU"},
        "roofline": dict(roofs[dom], kernel=KERNEL_NAMES[dom], family=dom),
        "kernel_rooflines": roofs,
        "kernel_ms": {f: round(fams[f][1], 3) for f in fams},
        "cpu_baseline": cpu,
    }
    out.update(legs)
    out.update(extra)
    if args.config == "c3" and not args.no_learned:
        # the learned stage's workload (2D GRF diffusion 1024^2) beside the headline
        out["learned_1024"] = learned_leg(args, sg, P, CONFIGS["c5"], False)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
