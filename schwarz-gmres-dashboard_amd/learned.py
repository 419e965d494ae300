"""Training data and training of the learned FGMRES stage's network.

Data comes from the plain two-level RAS-FGMRES solve itself.

* mode "krylov" (default, what the shipped weights use): the network runs on
  the c x c block averages of what the stage sees at run time.  For each
  seeded field (kappa = exp(g), g a Gaussian random field on a side x side
  grid, sigma 1.5, correlation length 0.05; b ~ N(0, 1) from the seed) a
  coefficient-weighted partition into `nparts` subdomains (overlap 2, FP32
  ILU(0), Nicolaides coarse space) drives FGMRES(30) from x0 = 0 for
  `cycles` restart cycles.  After the cycles listed in `sample_cycles` the
  stage's inputs are sampled: Arnoldi vectors v_j of the cycle (j in
  `columns`) and the residual r = b - A x it ended with; for each such vector
  u the input is s = u - A M u (the residual a two-level step leaves), as
  (<s>_c / rms(<s>_c), <g>_c) with <.>_c the c x c block average.  No target:
  loss kind 1 (the share of |s|^2 that the best multiple of A e fails to
  remove, A rebuilt from the averaged log kappa channel).
* mode "residual": the same without averaging, from restart-cycle residuals
  of the batched solve on a shared unweighted partition.
* mode "error": the solve continues to ||b - A x|| <= 1e-10 ||b|| for x*;
  input (r_k / rms(r_k), g), target e_k / rms(r_k) with e_k = x* - x_k, loss
  kind 0 (relative squared error).

Training: Adam (lr 1e-3) on batches of 16 samples with a cosine learning-rate
decay over 24 epochs, stopped after epoch 12 (the shipped weights: 1,400
fields, 16,800 samples); the residual loss is a poor guide to the solver (it
bottoms out after a few epochs while the iteration counts keep falling), so
the stopping epoch was chosen by solving held-out fields.

    python learned.py --out weights.npy
"""
import argparse
import importlib.util
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

CONVS = [(2, 16), (16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (64, 64), (64, 64),
         (128, 64), (64, 64), (96, 32), (32, 32), (48, 16), (16, 16)]


def init_params(seed):
    """He-normal weights, small biases, in the packed parameter order."""
    rng = np.random.default_rng(seed)
    p = []
    for cin, cout in CONVS:
        p.append(rng.standard_normal(cout * cin * 9) * np.sqrt(2.0 / (9 * cin)))
        p.append(np.zeros(cout))
    p.append(rng.standard_normal(16) * 0.1)
    p.append(np.zeros(1))
    return np.concatenate(p).astype(np.float32)


def load_problems():
    spec = importlib.util.spec_from_file_location("sgamd_problems", os.path.join(HERE, "problems.py"))
    P = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(P)
    return P


def generate_pairs(sg, ctx, P, nfields, side=256, base_seed=10000, per_field=10, restart=10,
                   sigma=1.5, corr=0.05, nparts=16, batch=32, log=None, mode="residual"):
    """Returns X [N, 2, side, side] and T [N, side, side] (float32, N = nfields *
    per_field); T is None in mode "residual"."""
    n = side * side
    rp0, ci0, _ = P.poisson2d(side, side)
    part1 = sg.partition(rp0, ci0, nparts).astype(np.int32)
    N = nfields * per_field
    X = np.empty((N, 2, side, side), np.float32)
    T = np.empty((N, side, side), np.float32) if mode == "error" else None
    k = 0
    t0 = time.perf_counter()
    for f0 in range(0, nfields, batch):
        seeds = list(range(base_seed + f0, base_seed + min(nfields, f0 + batch)))
        mats, gs, bs = [], [], []
        for s in seeds:
            g = P.grf(side, sigma, corr, s)
            mats.append(P.diffusion2d(np.exp(g)))
            gs.append(g)
            bs.append(P.normal(s + 7, n))
        ns = len(seeds)
        B = sg.BatchCsr(ctx, mats)
        part = np.concatenate([part1 + q * nparts for q in range(ns)]).astype(np.int32)
        M = sg.Precond(ctx, "twolevel_ras", nparts=nparts * ns, overlap=2, precision="fp32",
                       partition=part, coarse_solve=1).setup(B)
        b = np.concatenate(bs)
        db = sg.DeviceArray.from_host(ctx, b)
        dx = sg.DeviceArray(ctx, B.n)
        dx.zero()  # x0 = 0: device allocations are not initialised
        dr = sg.DeviceArray(ctx, B.n)
        dz = sg.DeviceArray(ctx, B.n)
        xs, rs = [], []
        for _ in range(per_field):
            sg.solve_batch(B, db, dx, M, method="fgmres", restart=restart, max_iter=1, tol=0.0,
                           ortho="cgs2")
            dr.upload(b)
            B.spmv(dx, dr, alpha=-1.0, beta=1.0)
            r = dr.get()
            if mode == "error":
                xs.append(dx.get())
                rs.append(r)
            else:  # s = r - A M r
                M.apply(dr, dz)
                B.spmv(dz, dr, alpha=-1.0, beta=1.0)
                rs.append(dr.get())
        xstar = None
        if mode == "error":
            res = sg.solve_batch(B, db, dx, M, method="fgmres", restart=30, max_iter=2000,
                                 rtol=1e-10, ortho="cgs2")
            assert all(r["converged"] for r in res)
            xstar = dx.get()
        for q in range(ns):
            sl = slice(q * n, (q + 1) * n)
            for j in range(per_field):
                r = rs[j][sl]
                rms = np.sqrt(np.mean(r * r))
                X[k, 0] = (r / rms).reshape(side, side)
                X[k, 1] = gs[q]
                if T is not None:
                    T[k] = ((xstar[sl] - xs[j][sl]) / rms).reshape(side, side)
                k += 1
        del M, B
        if log:
            log(f"# pairs: {k}/{N} ({time.perf_counter() - t0:.1f} s)")
    return X, T


def block_average(a, c):
    """c x c block averages of a 2-D array (sides multiples of c)."""
    h, w = a.shape
    return a.reshape(h // c, c, w // c, c).mean(axis=(1, 3)) if c > 1 else a


def generate_krylov_pairs(sg, ctx, P, nfields, side=1024, coarsen=4, base_seed=30000,
                          cycles=6, sample_cycles=(0, 2, 5), columns=(2, 12, 25), restart=30,
                          sigma=1.5, corr=0.05, nparts=256, log=None, threads=8):
    """Returns X [N, 2, side/coarsen, side/coarsen] float32 (mode "krylov")."""
    n = side * side
    cs = side // coarsen

    def field(seed):
        g = P.grf(side, sigma, corr, seed)
        rp, ci, v = P.diffusion2d(np.exp(g))
        part = sg.partition(rp, ci, nparts, 0, vals=v, weighted=True).astype(np.int32)
        return g, (rp, ci, v), part, P.normal(seed + 7, n)

    ds = sg.DeviceArray(ctx, n)
    dz = sg.DeviceArray(ctx, n)
    out = []

    def sample(A, M, u, lkc):
        ds.upload(u)
        M.apply(ds, dz)
        A.spmv(dz, ds, alpha=-1.0, beta=1.0)  # s = u - A M u
        sc = block_average(ds.get().reshape(side, side), coarsen)
        out.append(np.stack([sc / np.sqrt(np.mean(sc * sc)), lkc]).astype(np.float32))

    t0 = time.perf_counter()
    seeds = list(range(base_seed, base_seed + nfields))
    with ThreadPoolExecutor(threads) as ex:
        futs = [ex.submit(field, s) for s in seeds[:threads]]
        for i in range(nfields):
            g, mat, part, b = futs[i].result()
            futs[i] = None
            if i + threads < nfields:
                futs.append(ex.submit(field, seeds[i + threads]))
            A = sg.Csr(ctx, *mat)
            M = sg.Precond(ctx, "twolevel_ras", nparts=nparts, overlap=2, precision="fp32",
                           partition=part, coarse_solve=1).setup(A)
            db = sg.DeviceArray.from_host(ctx, b)
            dx = sg.DeviceArray(ctx, n)
            dx.zero()  # x0 = 0: device allocations are not initialised
            lkc = block_average(g, coarsen).astype(np.float32)
            for cyc in range(cycles):
                if cyc in sample_cycles:  # the residual this cycle starts from
                    dr = sg.DeviceArray.from_host(ctx, b)
                    A.spmv(dx, dr, alpha=-1.0, beta=1.0)
                    sample(A, M, dr.get(), lkc)
                sg.solve(A, db, dx, M, method="fgmres", restart=restart, max_iter=1, tol=0.0,
                         ortho="cgs2")
                if cyc in sample_cycles:
                    V = sg.last_basis(ctx, with_z=False)["V"]
                    for j in columns:
                        if j < V.shape[1]:
                            sample(A, M, V[:, j], lkc)
                    del V
            del M, A
            if log and (i % 50 == 0 or i == nfields - 1):
                log(f"# fields {i + 1}/{nfields}: {len(out)} samples "
                    f"({time.perf_counter() - t0:.1f} s)")
    return np.stack(out)


class _Ptr:
    def __init__(self, p):
        self.ptr = p


def train(sg, ctx, X, T, epochs=8, batch=16, lr=1e-3, seed=0, params=None, val=0.05, log=None,
          wd=0.0, keep_best=True, on_epoch=None, loss_kind=None):
    """Adam on device-resident data (loss kind 1 when T is None); returns
    (params, history): the parameters after the epoch with the lowest
    validation loss (keep_best, and a validation set) or after the last epoch.
    wd: decoupled weight decay of the Adam steps; on_epoch(epoch, net) runs
    after each epoch; loss_kind: 2 for the residual loss with the scale term
    (default: 0 with targets, 1 without)."""
    import ctypes as C
    N, _, H, W = X.shape
    nval = max(1, int(round(val * N))) if N > 1 else 0
    rng = np.random.default_rng(seed)
    perm = rng.permutation(N)
    X = X[perm]
    T = T[perm] if T is not None else None
    dX = sg.DeviceArray.from_host(ctx, X.reshape(-1))
    dT = sg.DeviceArray.from_host(ctx, T.reshape(-1)) if T is not None else None
    net = sg.UNet(ctx, init_params(seed) if params is None else params)
    net.set_loss(loss_kind if loss_kind is not None else (0 if T is not None else 1))
    net.set_weight_decay(wd)
    xs, ts = 2 * H * W * 4, H * W * 4

    def views(i, b):
        i = int(i)
        vt = _Ptr(C.c_void_p(dT.ptr.value + i * ts)) if dT is not None else None
        return _Ptr(C.c_void_p(dX.ptr.value + i * xs)), vt

    ntr = N - nval
    nb = max(1, ntr // batch)
    total = epochs * nb
    hist = []
    step = 0
    best, best_val = None, np.inf
    t0 = time.perf_counter()
    for ep in range(epochs):
        order = rng.permutation(nb)
        tl = 0.0
        for j in order:
            vx, vt = views(j * batch, batch)
            tl += net.loss_grad_dev(batch, H, W, vx, vt)
            a = lr * 0.5 * (1.0 + np.cos(np.pi * step / total))
            net.adam(max(a, 1e-3 * lr))
            step += 1
        vl = 0.0
        if nval:
            for i in range(ntr, N, batch):
                bb = min(batch, N - i)
                vx, vt = views(i, bb)
                vl += net.loss_grad_dev(bb, H, W, vx, vt) * bb
            vl /= nval
        hist.append({"epoch": ep + 1, "train_loss": tl / nb, "val_loss": vl,
                     "s": round(time.perf_counter() - t0, 2)})
        if keep_best and nval and vl < best_val:
            best, best_val = net.get_params(), vl
        if on_epoch:
            on_epoch(ep + 1, net)
        if log:
            log(f"# epoch {ep + 1}: train {tl / nb:.4f} val {vl:.4f} "
                f"({time.perf_counter() - t0:.1f} s)")
    return (best if best is not None else net.get_params()), hist


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--fields", type=int, default=1400)
    ap.add_argument("--side", type=int, default=1024)
    ap.add_argument("--coarsen", type=int, default=4)
    ap.add_argument("--per-field", type=int, default=10)
    ap.add_argument("--epochs", type=int, default=24, help="length of the cosine schedule")
    ap.add_argument("--stop-epoch", type=int, default=12,
                    help="keep the parameters after this epoch and stop (0: run all)")
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--mode", default="krylov", choices=("krylov", "residual", "error"))
    ap.add_argument("--out", required=True)
    args = ap.parse_args()
    sys.path.insert(0, os.path.dirname(HERE))
    import __graft_entry__
    sg = __graft_entry__.load_package()
    P = load_problems()
    ctx = sg.Context(0)
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    t0 = time.perf_counter()
    if args.mode == "krylov":
        X, T = generate_krylov_pairs(sg, ctx, P, args.fields, side=args.side,
                                     coarsen=args.coarsen, log=log), None
    else:
        X, T = generate_pairs(sg, ctx, P, args.fields, side=args.side, per_field=args.per_field,
                              log=log, mode=args.mode)
    tgen = time.perf_counter() - t0

    class _Stop(Exception):
        pass

    kept = {}

    def at_epoch(ep, net):
        if args.stop_epoch and ep >= args.stop_epoch:
            kept["params"] = net.get_params()
            raise _Stop()

    hist = None
    try:
        params, hist = train(sg, ctx, X, T, epochs=args.epochs, batch=args.batch, lr=args.lr,
                             seed=args.seed, log=log, keep_best=False, on_epoch=at_epoch)
    except _Stop:
        params = kept["params"]
    np.save(args.out, params)
    print(json.dumps({"pairs": int(X.shape[0]), "generate_s": round(tgen, 1),
                      "epochs_run": args.stop_epoch or args.epochs,
                      "history": hist}))


if __name__ == "__main__":
    main()
