"""The learned FGMRES stage (sg_precond_create_learned):

* fallback: when no network correction is accepted the stage is exactly the
  base two-level Schwarz preconditioner, so the FGMRES solve is bitwise the
  plain one;
* an accepted correction never increases the base step's residual;
* it needs the flexible method;
* a network trained on pairs from the plain solve (tiny run) is used by the
  solve, which still reaches an FP64-verified 1e-8.
"""
import importlib.util
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sg():
    import __graft_entry__
    return __graft_entry__.load_package()


@pytest.fixture(scope="module")
def mods():
    import __graft_entry__
    out = []
    for name in ("problems", "learned"):
        spec = importlib.util.spec_from_file_location(
            "sgamd_" + name, os.path.join(__graft_entry__.PKG_DIR, name + ".py"))
        m = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(m)
        out.append(m)
    return out


@pytest.fixture(scope="module")
def ctx(sg):
    c = sg.Context(0)
    yield c
    c.close()


def dev(sg, ctx, a):
    return sg.DeviceArray.from_host(ctx, np.ascontiguousarray(a, np.float64))


def grf_problem(P, side, seed):
    g = P.grf(side, 1.5, 0.05, seed)
    rp, ci, v = P.diffusion2d(np.exp(g))
    return rp, ci, v, g, P.normal(seed + 7, side * side)


def test_fallback_is_the_plain_solve(sg, ctx, mods):
    P, Lm = mods
    side = 64
    rp, ci, v, g, b = grf_problem(P, side, 21)
    A = sg.Csr(ctx, rp, ci, v)
    base = sg.Precond(ctx, "twolevel_ras", nparts=8, overlap=2, precision="fp32").setup(A)
    net = sg.UNet(ctx, Lm.init_params(3))
    stage = sg.LearnedStage(ctx, base, net, g.ravel(), side, side, min_gain=1.0001).setup(A)
    kw = dict(method="fgmres", restart=20, max_iter=500, rtol=1e-8, ortho="cgs2")
    x1, x2 = dev(sg, ctx, np.zeros(A.n)), dev(sg, ctx, np.zeros(A.n))
    db = dev(sg, ctx, b)
    r1 = sg.solve(A, db, x1, base, **kw)
    r2 = sg.solve(A, db, x2, stage, **kw)
    assert r1["converged"] and r2["converged"]
    assert r1["iterations"] == r2["iterations"]
    assert np.array_equal(x1.get(), x2.get())
    calls, acc = stage.counts()
    assert acc == 0 and calls == r2["precond_applies"] > 0


def test_accepted_correction_reduces_the_residual(sg, ctx, mods):
    P, Lm = mods
    side = 32
    rp, ci, v, g, _ = grf_problem(P, side, 22)
    A = sg.Csr(ctx, rp, ci, v)
    base = sg.Precond(ctx, "twolevel_ras", nparts=4, overlap=2, precision="fp32").setup(A)
    net = sg.UNet(ctx, Lm.init_params(4))
    stage = sg.LearnedStage(ctx, base, net, g.ravel(), side, side, min_gain=1e-12).setup(A)
    rng = np.random.default_rng(5)
    from scipy.sparse import csr_matrix
    As = csr_matrix((v, ci, rp), shape=(A.n, A.n))
    for trial in range(3):
        x = rng.standard_normal(A.n)
        zb, z = dev(sg, ctx, np.zeros(A.n)), dev(sg, ctx, np.zeros(A.n))
        base.apply(dev(sg, ctx, x), zb)
        stage.apply(dev(sg, ctx, x), z)
        rb = np.linalg.norm(x - As @ zb.get())
        rz = np.linalg.norm(x - As @ z.get())
        assert rz <= rb * (1 + 1e-12)
    calls, acc = stage.counts()
    assert calls == 3 and acc == 3


def _pool(a, c):
    h, w = a.shape
    return a.reshape(h // c, c, w // c, c).mean(axis=(1, 3))


def _interp(o, c):
    """Cell-centred bilinear interpolation of the H x W grid o onto cH x cW,
    clamped at the edges (the identity for c = 1)."""
    H, W = o.shape

    def axis(m, k):
        f = (np.arange(c * m) + 0.5) / c - 0.5
        i0 = np.floor(f).astype(int)
        t = f - i0
        return np.clip(i0, 0, m - 1), np.clip(i0 + 1, 0, m - 1), t

    x0, x1, tx = axis(W, c)
    y0, y1, ty = axis(H, c)
    top = (1 - tx)[None, :] * o[y0][:, x0] + tx[None, :] * o[y0][:, x1]
    bot = (1 - tx)[None, :] * o[y1][:, x0] + tx[None, :] * o[y1][:, x1]
    return (1 - ty)[:, None] * top + ty[:, None] * bot


def grf_grid(P, nx, ny, seed):
    """A GRF diffusion problem on an nx x ny grid (a window of a square field)."""
    g = P.grf(max(nx, ny), 1.5, 0.05, seed)[:ny, :nx]
    rp, ci, v = P.diffusion2d(np.exp(g))
    return rp, ci, v, g


# (nx, ny): 64 x 64 and 128 x 96 run the tiled A e kernel on the SELL-64 copy,
# 96 x 64 the generic row kernels (nx not a multiple of 64)
@pytest.mark.parametrize("nx,ny", [(64, 64), (128, 96), (96, 64)])
@pytest.mark.parametrize("coarsen", [1, 2, 4])
def test_stage_matches_its_definition(sg, ctx, mods, coarsen, nx, ny):
    """One application against a numpy restatement of the stage: s = x - A z_b,
    its c x c block averages sc normalised by their rms (plus the averaged
    log kappa) into the network, e = rms * bilinear(net output), z = z_b +
    (s.Ae / Ae.Ae) e.  The network output is taken from the same network's
    own forward on the restated input."""
    from scipy.sparse import csr_matrix
    P, Lm = mods
    rp, ci, v, g = grf_grid(P, nx, ny, 24)
    A = sg.Csr(ctx, rp, ci, v)
    As = csr_matrix((v, ci, rp), shape=(A.n, A.n))
    base = sg.Precond(ctx, "twolevel_ras", nparts=4, overlap=2, precision="fp32").setup(A)
    net = sg.UNet(ctx, Lm.init_params(6))
    stage = sg.LearnedStage(ctx, base, net, g.ravel(), nx, ny, min_gain=1e-12,
                            coarsen=coarsen).setup(A)
    rng = np.random.default_rng(11 + coarsen)
    for trial in range(2):
        x = rng.standard_normal(A.n)
        zb, z = dev(sg, ctx, np.zeros(A.n)), dev(sg, ctx, np.zeros(A.n))
        base.apply(dev(sg, ctx, x), zb)
        stage.apply(dev(sg, ctx, x), z)
        zb = zb.get()
        s = x - As @ zb
        sc = _pool(s.reshape(ny, nx), coarsen)
        q = float((sc * sc).sum())
        rms = np.sqrt(q / sc.size)
        nin = np.stack([sc / rms, _pool(g, coarsen)])[None].astype(np.float32)
        o = net.forward(nin)[0].astype(np.float64)
        e = rms * _interp(o, coarsen).ravel()
        Ae = As @ e
        alpha = (s @ Ae) / (Ae @ Ae)
        corr = z.get() - zb
        assert np.linalg.norm(corr - alpha * e) <= 1e-4 * np.linalg.norm(alpha * e), trial
    calls, acc = stage.counts()
    assert calls == 2 and acc == 2


@pytest.mark.parametrize("nx,ny", [(128, 96), (96, 64)])
def test_stage_hands_fgmres_a_consistent_product(sg, ctx, mods, nx, ny):
    """Inside FGMRES the stage returns w = A z by linearity (A z_b + alpha A e)
    instead of the step's SpMV: the last cycle's basis must still satisfy the
    Arnoldi relation A Z = V Hbar to rounding."""
    from scipy.sparse import csr_matrix
    P, Lm = mods
    rp, ci, v, g = grf_grid(P, nx, ny, 25)
    A = sg.Csr(ctx, rp, ci, v)
    As = csr_matrix((v, ci, rp), shape=(A.n, A.n))
    base = sg.Precond(ctx, "twolevel_ras", nparts=4, overlap=2, precision="fp32").setup(A)
    net = sg.UNet(ctx, Lm.init_params(7))
    stage = sg.LearnedStage(ctx, base, net, g.ravel(), nx, ny, min_gain=1e-12, coarsen=4).setup(A)
    b = np.random.default_rng(3).standard_normal(A.n)
    x = dev(sg, ctx, np.zeros(A.n))
    r = sg.solve(A, dev(sg, ctx, b), x, stage, method="fgmres", restart=12, max_iter=1,
                 tol=1e-300, ortho="cgs2")
    B = sg.last_basis(ctx)
    k, kv = B["k"], B["kv"]
    assert r["iterations"] == 12 and k == 12 and kv == 12  # one full cycle
    V, Z, H = B["V"], B["Z"], B["Hbar"]
    AZ = As @ Z
    # |A|_2 <= sqrt(|A|_1 |A|_inf)
    Aabs = abs(As)
    norm2 = np.sqrt(Aabs.sum(axis=0).max() * Aabs.sum(axis=1).max())
    for j in range(kv - 1):
        res = AZ[:, j] - V[:, :j + 2] @ H[:j + 2, j]
        assert np.linalg.norm(res) <= 1e-12 * norm2 * np.linalg.norm(Z[:, j]), j
    calls, acc = stage.counts()
    assert calls == k and acc > 0


def test_needs_fgmres(sg, ctx, mods):
    P, Lm = mods
    rp, ci, v, g, b = grf_problem(P, 16, 23)
    A = sg.Csr(ctx, rp, ci, v)
    base = sg.Precond(ctx, "ras", nparts=2, overlap=1).setup(A)
    net = sg.UNet(ctx, Lm.init_params(5))
    stage = sg.LearnedStage(ctx, base, net, g.ravel(), 16, 16).setup(A)
    with pytest.raises(sg.SgError):
        sg.solve(A, dev(sg, ctx, b), dev(sg, ctx, np.zeros(A.n)), stage, method="gmres")


def test_tiny_training_run_drives_the_stage(sg, ctx, mods):
    P, Lm = mods
    X, T = Lm.generate_pairs(sg, ctx, P, nfields=8, side=32, base_seed=500, per_field=4,
                             restart=4, nparts=4, batch=4, mode="error")
    assert X.shape == (32, 2, 32, 32) and np.isfinite(X).all() and np.isfinite(T).all()
    Xr, Tr = Lm.generate_pairs(sg, ctx, P, nfields=8, side=32, base_seed=500, per_field=4,
                               restart=4, nparts=4, batch=4, mode="residual")
    assert Tr is None and np.isfinite(Xr).all()
    _, hr = Lm.train(sg, ctx, Xr, None, epochs=20, batch=4, lr=2e-3, seed=2)
    assert hr[-1]["train_loss"] < hr[0]["train_loss"]
    params, hist = Lm.train(sg, ctx, X, T, epochs=30, batch=4, lr=2e-3, seed=1)
    assert hist[-1]["train_loss"] < hist[0]["train_loss"]
    side = 32
    rp, ci, v, g, b = grf_problem(P, side, 777)
    A = sg.Csr(ctx, rp, ci, v)
    base = sg.Precond(ctx, "twolevel_ras", nparts=4, overlap=2, precision="fp32").setup(A)
    net = sg.UNet(ctx, params)
    stage = sg.LearnedStage(ctx, base, net, g.ravel(), side, side).setup(A)
    x = dev(sg, ctx, np.zeros(A.n))
    r = sg.solve(A, dev(sg, ctx, b), x, stage, method="fgmres", restart=20, max_iter=500,
                 rtol=1e-8, ortho="cgs2")
    assert r["converged"] and r["rel_true_residual"] <= 1e-8
    calls, acc = stage.counts()
    assert calls == r["precond_applies"]


def test_shipped_network_gains_on_held_out_fields(sg, ctx, mods):
    """Ablation with the shipped weights (trained on samples from the plain
    solves of fields 30000..30999) on held-out 1024^2 fields (seeds 20000..
    20002, sigma 1.5, correlation length 0.05), with the 1024^2 workload's
    solver: mixed two-level RAS(256 subdomains, overlap 2, coefficient-weighted
    partition)-FGMRES(30) to an FP64-verified 1e-8, with and without the
    learned stage (network on 4 x 4 block averages).  The stage must cut the
    total iteration count by more than 40% and the total time; both are
    printed.  Each solve is timed after an untimed warm-up solve."""
    import time
    P, Lm = mods
    import __graft_entry__
    params = np.load(os.path.join(__graft_entry__.PKG_DIR, "weights", "unet_grf1024_c4.npy"))
    side = 1024
    net = sg.UNet(ctx, params)
    kw = dict(method="fgmres", restart=30, max_iter=3000, rtol=1e-8, ortho="cgs2")
    it_plain = it_learn = 0
    t_plain = t_learn = 0.0
    for seed in range(20000, 20003):
        rp, ci, v, g, b = grf_problem(P, side, seed)
        A = sg.Csr(ctx, rp, ci, v)
        part = sg.partition(rp, ci, 256, 0, vals=v, weighted=True)
        base = sg.Precond(ctx, "twolevel_ras", nparts=256, overlap=2, precision="fp32",
                          partition=part, coarse_solve=1).setup(A)
        stage = sg.LearnedStage(ctx, base, net, g.ravel(), side, side, coarsen=4).setup(A)
        db = dev(sg, ctx, b)
        x = dev(sg, ctx, np.zeros(A.n))
        for M, which in ((base, "plain"), (stage, "learned")):
            for timed in (False, True):
                x.zero()
                ctx.sync()
                t0 = time.perf_counter()
                r = sg.solve(A, db, x, M, **kw)
                dt = time.perf_counter() - t0
            assert r["converged"] and r["rel_true_residual"] <= 1e-8, (seed, which, r)
            if which == "plain":
                it_plain += r["iterations"]
                t_plain += dt
            else:
                it_learn += r["iterations"]
                t_learn += dt
    gain = 1.0 - it_learn / it_plain
    print(f"held-out 1024^2 fields: iterations {it_plain} plain, {it_learn} with the learned "
          f"stage ({100 * gain:.1f}% fewer); time {t_plain:.3f} s vs {t_learn:.3f} s "
          f"({t_plain / t_learn:.2f}x)")
    assert it_learn < 0.6 * it_plain
    assert t_learn < t_plain
