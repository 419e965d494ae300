"""sg_solve_noalloc (the no-allocation variant of sg_solve, whose workspace is
a static device array) against sg_solve, through the C ABI: two calls in a
row, the second on a smaller system, each equal to sg_solve bit for bit
(iterations, cycles, x, residual history, true residual); sizes beyond the
static array and batch matrices are rejected before anything runs."""
import ctypes as C
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
def P():
    import __graft_entry__
    spec = importlib.util.spec_from_file_location(
        "sgamd_problems", os.path.join(__graft_entry__.PKG_DIR, "problems.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def _noalloc(sg, A, b, x, M, method, restart, max_iter, tol=0.0, rtol=0.0, ortho="cgs2"):
    p = sg.SolveParams(sg.METHODS[method], restart, max_iter, tol, rtol,
                       1 if ortho == "cgs2" else 0, 0, 0)
    r = sg.SolveResult()
    cap = max_iter + restart + 2
    hist = np.zeros(cap)
    st = sg.lib().sg_solve_noalloc(A.h, M.h if M is not None else None, b.ptr, x.ptr,
                                   C.byref(p), C.byref(r), hist.ctypes.data_as(C.c_void_p), cap)
    return st, sg._result(r, hist, cap)


def _compare(sg, ctx, rp, ci, v, b, M_kw, kw):
    A = sg.Csr(ctx, rp, ci, v)
    M = sg.Precond(ctx, **M_kw).setup(A) if M_kw else None
    n = len(rp) - 1
    db = sg.DeviceArray.from_host(ctx, b)
    x1 = sg.DeviceArray(ctx, n)
    x1.zero()
    r1 = sg.solve(A, db, x1, M, **kw)
    x2 = sg.DeviceArray(ctx, n)
    x2.zero()
    st, r2 = _noalloc(sg, A, db, x2, M, kw["method"], kw["restart"], kw["max_iter"],
                      kw.get("tol", 0.0), kw.get("rtol", 0.0), kw.get("ortho", "cgs2"))
    assert st == 0, sg.lib().sg_last_error(ctx.h).decode()
    for key in ("iterations", "converged", "cycles", "precond_applies"):
        assert r1[key] == r2[key], (key, r1[key], r2[key])
    assert np.array_equal(r1["hist"], r2["hist"])
    assert r1["residual_norm"] == r2["residual_norm"]
    assert r1["rel_true_residual"] == r2["rel_true_residual"]
    assert np.array_equal(x1.get(), x2.get())
    return r1


def test_noalloc_equals_sg_solve_twice_shrinking(sg, P):
    ctx = sg.Context(0)
    # first call: 64^2 GRF diffusion, FGMRES(30) + two-level RAS FP32, CGS2, rtol
    rp, ci, v = P.diffusion2d(np.exp(P.grf(64, 1.0, 0.1, 3)))
    b = P.normal(5, len(rp) - 1)
    part = sg.partition(rp, ci, 8, 0)
    r = _compare(sg, ctx, rp, ci, v, b,
                 dict(kind="twolevel_ras", nparts=8, overlap=1, precision="fp32",
                      partition=part, coarse_solve=1),
                 dict(method="fgmres", restart=30, max_iter=500, rtol=1e-8, ortho="cgs2"))
    assert r["converged"] and r["iterations"] > 10
    # second call, smaller system: 24^2 Poisson, left GMRES(20) + ILU(0), MGS,
    # absolute tolerance (the reference's semantics)
    rp, ci, v = P.poisson2d(24, 24)
    xs = P.uniform_pm1(1, len(rp) - 1)
    b = np.zeros(len(rp) - 1)
    for i in range(len(rp) - 1):
        b[i] = np.dot(v[rp[i]:rp[i + 1]], xs[ci[rp[i]:rp[i + 1]]])
    r = _compare(sg, ctx, rp, ci, v, b, dict(kind="ilu0"),
                 dict(method="gmres", restart=20, max_iter=400, tol=1e-10, ortho="mgs"))
    assert r["converged"]
    # no preconditioner, capped: the iteration bound ends it unconverged
    r = _compare(sg, ctx, rp, ci, v, b, None,
                 dict(method="fgmres", restart=7, max_iter=9, tol=1e-14, ortho="cgs2"))
    assert not r["converged"] and r["iterations"] == 14
    ctx.close()


def test_noalloc_rejects_what_it_cannot_hold(sg, P):
    ctx = sg.Context(0)
    rp, ci, v = P.poisson2d(1024, 1024)  # FGMRES(30) basis alone: 520 MB
    A = sg.Csr(ctx, rp, ci, v)
    n = len(rp) - 1
    db, dx = sg.DeviceArray(ctx, n), sg.DeviceArray(ctx, n)
    db.zero()
    dx.zero()
    st, _ = _noalloc(sg, A, db, dx, None, "fgmres", 30, 10, tol=1e-8)
    assert st == 1 and "scratch" in sg.lib().sg_last_error(ctx.h).decode()
    systems = [P.poisson2d(4, 4), P.poisson2d(3, 5)]
    B = sg.BatchCsr(ctx, systems)
    db, dx = sg.DeviceArray(ctx, B.n), sg.DeviceArray(ctx, B.n)
    db.zero()
    dx.zero()
    st, _ = _noalloc(sg, B, db, dx, None, "gmres", 5, 10, tol=1e-8)
    assert st == 1
    ctx.close()
