"""Batched many-system solve (sg_solve_batch): every system of the batch must
reproduce its own single-system solve bit for bit -- same iteration count,
same cycles, same x -- because the batch runs the same kernels over the same
per-system reduction ranges; the captured-graph cycle must equal the directly
launched one."""
import importlib.util
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.join(os.path.dirname(__file__), "..")


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


@pytest.fixture(scope="module")
def ctx(sg):
    c = sg.Context(0)
    yield c
    c.close()


def dev(sg, ctx, a):
    return sg.DeviceArray.from_host(ctx, np.ascontiguousarray(a, np.float64))


def _small_batch(P, nsys=6, seed=77):
    systems = P.batch_systems(nsys, base_seed=seed, side_range=(12, 40))
    systems.append(P.poisson2d(3, 1) + (np.array([1.0, -2.0, 0.5]),))  # a 3-row system
    return systems


def _solve_each(sg, ctx, systems, parts, kind, precision, kw, graph=False):
    out = []
    for (rp, ci, v, b), part, k in zip(systems, parts, [p.max() + 1 for p in parts]):
        A = sg.Csr(ctx, rp, ci, v)
        M = sg.Precond(ctx, kind, nparts=int(k), overlap=2, precision=precision,
                       partition=part).setup(A)
        x = dev(sg, ctx, np.zeros(len(b)))
        r = sg.solve(A, dev(sg, ctx, b), x, M, graph=graph, **kw)
        out.append((r, x.get()))
    return out


def _batch(sg, ctx, systems, kind, precision, kw, graph=True, nparts=None):
    mats = [(rp, ci, v) for rp, ci, v, _ in systems]
    nparts = nparts or [min(3, max(1, (len(rp) - 1) // 200)) for rp, _, _ in mats]
    part = sg.batch_partition(mats, nparts)
    B = sg.BatchCsr(ctx, mats)
    M = sg.Precond(ctx, kind, nparts=int(sum(nparts)), overlap=2, precision=precision,
                   partition=part).setup(B)
    b = np.concatenate([s[3] for s in systems])
    x = dev(sg, ctx, np.zeros(B.n))
    res = sg.solve_batch(B, dev(sg, ctx, b), x, M, graph=graph, **kw)
    xs = np.split(x.get(), B.sys_rows[1:-1])
    parts = []
    for q in range(len(mats)):
        p = part[B.sys_rows[q]:B.sys_rows[q + 1]]
        parts.append(p - p.min())
    return res, xs, parts


@pytest.mark.parametrize("kind,precision,method,ortho", [
    ("ras", "fp32", "fgmres", "cgs2"),
    ("ras", "fp64", "gmres", "mgs"),
    ("asm", "fp64", "fgmres", "cgs2"),
    # block-diagonal coarse matrix: the LU coarse solve factors it block by block
    ("twolevel_ras", "fp64", "fgmres", "cgs2"),
])
def test_batch_equals_single_solves(sg, ctx, P, kind, precision, method, ortho):
    systems = _small_batch(P)
    kw = dict(method=method, restart=12, max_iter=400, rtol=1e-8, ortho=ortho)
    res, xs, parts = _batch(sg, ctx, systems, kind, precision, kw)
    single = _solve_each(sg, ctx, systems, parts, kind, precision, kw)
    iters = set()
    for q, ((r1, x1), rb, xb) in enumerate(zip(single, res, xs)):
        assert rb["converged"] and r1["converged"], q
        assert rb["iterations"] == r1["iterations"], q
        assert rb["cycles"] == r1["cycles"], q
        assert rb["precond_applies"] == r1["precond_applies"], q
        assert np.array_equal(xb, x1), q
        assert np.array_equal(rb["hist"], r1["hist"]), q
        assert rb["rel_true_residual"] <= 1e-8
        iters.add(rb["iterations"])
    assert len(iters) > 2  # systems finish at different steps


def test_graph_equals_direct_launches(sg, ctx, P):
    systems = _small_batch(P, nsys=4, seed=91)
    kw = dict(method="fgmres", restart=10, max_iter=300, rtol=1e-9, ortho="cgs2")
    rg, xg, _ = _batch(sg, ctx, systems, "ras", "fp32", kw, graph=True)
    rd, xd, _ = _batch(sg, ctx, systems, "ras", "fp32", kw, graph=False)
    for a, b, xa, xb in zip(rg, rd, xg, xd):
        assert a["iterations"] == b["iterations"]
        assert np.array_equal(xa, xb)
    # single-system solve through the captured graph equals the direct one
    rp, ci, v, b = systems[0]
    A = sg.Csr(ctx, rp, ci, v)
    M = sg.Precond(ctx, "ras", nparts=3, overlap=1, precision="fp32").setup(A)
    out = []
    for graph in (False, True, True):  # a second graph solve replays from scratch
        x = dev(sg, ctx, np.zeros(len(b)))
        r = sg.solve(A, dev(sg, ctx, b), x, M, graph=graph, **kw)
        out.append((r["iterations"], x.get()))
    for it, x in out[1:]:
        assert it == out[0][0] and np.array_equal(x, out[0][1])


def test_batch_absolute_tolerance_and_caps(sg, ctx, P):
    """Reference semantics per system: absolute tol, max_iter bounding cycles."""
    systems = _small_batch(P, nsys=3, seed=5)
    kw = dict(method="gmres", restart=8, max_iter=3, tol=1e-30, ortho="cgs2")
    res, xs, parts = _batch(sg, ctx, systems, "ras", "fp64", kw)
    single = _solve_each(sg, ctx, systems, parts, "ras", "fp64", kw)
    for (r1, x1), rb, xb in zip(single, res, xs):
        assert not rb["converged"] and rb["cycles"] == r1["cycles"]
        assert rb["iterations"] == r1["iterations"]
        assert np.array_equal(xb, x1)


def test_batch_errors(sg, ctx, P):
    systems = _small_batch(P, nsys=2, seed=3)
    mats = [(rp, ci, v) for rp, ci, v, _ in systems]
    B = sg.BatchCsr(ctx, mats)
    b = dev(sg, ctx, np.ones(B.n))
    x = dev(sg, ctx, np.zeros(B.n))
    with pytest.raises(sg.SgError):  # a batch matrix goes through sg_solve_batch
        sg.solve(B, b, x, None)
    rp, ci, v = mats[0]
    n = len(rp) - 1
    # coupling across the system boundary is rejected
    bad_ci = np.array(ci, copy=True)
    bad_ci[-1] = n  # last row of system 0 points into system 1
    with pytest.raises(sg.SgError):
        sg.BatchCsr(ctx, [(rp, bad_ci, v), mats[1]])


def test_batch_zero_rhs_system(sg, ctx, P):
    """A system of the batch with b = 0: in true-residual mode it is converged
    at the first restart (x0 = 0 is exact) instead of cycling to max_iter, and
    the other systems still equal their single solves."""
    systems = _small_batch(P, nsys=3, seed=13)
    rp, ci, v, b = systems[1]
    systems[1] = (rp, ci, v, np.zeros_like(b))
    kw = dict(method="fgmres", restart=10, max_iter=300, rtol=1e-8, ortho="cgs2")
    res, xs, parts = _batch(sg, ctx, systems, "ras", "fp32", kw)
    single = _solve_each(sg, ctx, systems, parts, "ras", "fp32", kw)
    assert res[1]["converged"] and res[1]["iterations"] == 0 and res[1]["cycles"] == 1
    assert not np.any(xs[1])
    for q, ((r1, x1), rb, xb) in enumerate(zip(single, res, xs)):
        assert rb["converged"] and r1["converged"], q
        assert rb["iterations"] == r1["iterations"], q
        assert np.array_equal(xb, x1), q


def test_batch_offsets_validated_before_use(sg, ctx):
    """sg_csr_create_batch rejects system offsets that do not start at 0, run
    past the rows or decrease, before it reads row_ptr at them."""
    import ctypes as C
    L = sg.lib()
    rp = np.array([0, 1, 2, 3, 4], np.int32)
    ci = np.array([0, 1, 2, 3], np.int32)
    v = np.ones(4)
    for off in ([-1000000, 2, 4], [0, 9999999, 4], [0, 3, 2, 4], [1, 2, 4]):
        o = np.array(off, np.int32)
        h = C.c_void_p()
        st = L.sg_csr_create_batch(ctx.h, len(off) - 1, o.ctypes.data_as(C.c_void_p), 4,
                                   rp.ctypes.data_as(C.c_void_p), ci.ctypes.data_as(C.c_void_p),
                                   v.ctypes.data_as(C.c_void_p), C.byref(h))
        assert st == 1 and not h.value, off  # SG_E_ARG
