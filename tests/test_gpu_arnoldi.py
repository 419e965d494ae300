"""The Arnoldi process of the device solver, checked on the basis it actually
built (sg_solve_last_basis), against the mathematics that fixes it:

* orthogonality:  ||V^T V - I||_F <= 1e-12 for the Gram-corrected CGS2 that
  every mixed-precision solve runs (two passes over V, re-orthogonalisation
  coefficients from the lagged Gram matrix, h = (2I - G) p);
* the Arnoldi relation of the reference's loops, column by column:
  FGMRES  A z_j = V_{j+2} Hbar[:j+2, j]            (src/core/fgmres.cpp:77-107)
  GMRES   M^-1 A v_j = V_{j+2} Hbar[:j+2, j]        (src/core/gmres.cpp:84-128)
  to 1e-12 ||A||_2 ||z_j|| (resp. ||M^-1 A v_j||);
* the same for MGS in the reference's order as the control (its orthogonality
  is reported, not held to 1e-12: MGS loses it in proportion to the
  conditioning of the basis).

Matrix: 24^3 finite-volume diffusion with kappa = 10^u, u in {-3, 0, 3} on 8^3
blocks (face contrast up to 1e6), b ~ N(0, 1); one restart cycle of 30 steps.
"""
import importlib.util
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sg():
    import __graft_entry__
    return __graft_entry__.load_package()


@pytest.fixture(scope="module")
def prob():
    import __graft_entry__
    spec = importlib.util.spec_from_file_location(
        "sgamd_problems", os.path.join(__graft_entry__.PKG_DIR, "problems.py"))
    P = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(P)
    rp, ci, v = P.diffusion3d(P.block_jump_kappa(24, 8, 11))
    n = len(rp) - 1
    A = sp.csr_matrix((v, ci, rp), shape=(n, n))
    norm2 = float(spla.eigsh(A, k=1, which="LA", return_eigenvectors=False)[0])
    return dict(rp=rp, ci=ci, v=v, A=A, n=n, b=P.normal(21, n), norm2=norm2,
                part=None)


@pytest.fixture(scope="module")
def ctx(sg):
    c = sg.Context(0)
    yield c
    c.close()


def _run(sg, ctx, prob, kind, precision, method, ortho, m=30):
    A = sg.Csr(ctx, prob["rp"], prob["ci"], prob["v"])
    M = None
    if kind != "none":
        part = sg.partition(prob["rp"], prob["ci"], 16, 300, vals=prob["v"], weighted=True)
        M = sg.Precond(ctx, kind, nparts=16, overlap=1, precision=precision, partition=part,
                       coarse_solve=1).setup(A)
    db = sg.DeviceArray.from_host(ctx, prob["b"])
    dx = sg.DeviceArray(ctx, prob["n"])
    dx.zero()
    r = sg.solve(A, db, dx, M, method=method, restart=m, max_iter=1, tol=1e-300, ortho=ortho)
    B = sg.last_basis(ctx)
    assert B["m"] == m and B["n"] == prob["n"]
    assert r["iterations"] == m and B["k"] == m  # one full cycle, no breakdown
    assert B["kv"] == m  # the m-th step forms no v_m
    return A, M, B


def _orth(V):
    return np.linalg.norm(V.T @ V - np.eye(V.shape[1]))


def _check_relation(Y, V, H, norm2, tol=1e-12):
    """Column j of Y against V[:, :j+2] Hbar[:j+2, j] for j < kv - 1."""
    kv = V.shape[1]
    worst = 0.0
    for j in range(kv - 1):
        res = Y[:, j] - V[:, :j + 2] @ H[:j + 2, j]
        rel = np.linalg.norm(res) / (norm2 * np.linalg.norm(Y[:, j]))
        worst = max(worst, rel)
    assert worst <= tol, worst
    return worst


@pytest.mark.parametrize("kind,precision", [("twolevel_ras", "fp32"), ("ras", "fp64"),
                                            ("none", "fp64")])
@pytest.mark.parametrize("ortho", ["cgs2", "mgs"])
def test_fgmres_arnoldi_relation_and_orthogonality(sg, ctx, prob, kind, precision, ortho):
    _, _, B = _run(sg, ctx, prob, kind, precision, "fgmres", ortho)
    V, H = B["V"], B["Hbar"]
    Z = B["Z"] if kind != "none" else V[:, :B["k"]]
    AZ = prob["A"] @ Z
    rel = _check_relation(AZ, V, H, prob["norm2"])
    orth = _orth(V)
    print(f"fgmres {kind}/{precision} {ortho}: ||V^T V - I|| = {orth:.2e}, "
          f"relation {rel:.2e}")
    if ortho == "cgs2":
        assert orth <= 1e-12, orth
    else:
        assert orth <= 1e-6, orth
    # the first basis vector is r0 / ||r0|| with x0 = 0
    b = prob["b"]
    assert np.allclose(V[:, 0], b / np.linalg.norm(b), rtol=0, atol=1e-15)
    assert abs(np.linalg.norm(V[:, 0]) - 1.0) <= 1e-14


@pytest.mark.parametrize("ortho", ["cgs2", "mgs"])
def test_left_gmres_arnoldi_relation(sg, ctx, prob, ortho):
    A, M, B = _run(sg, ctx, prob, "ras", "fp64", "gmres", ortho)
    V, H = B["V"], B["Hbar"]
    n, kv = prob["n"], B["kv"]
    # M^-1 A v_j with the device's own SpMV and preconditioner (FP64 RAS is a
    # fixed linear operator, so these are the vectors the solve produced)
    Y = np.zeros((n, kv - 1))
    dv, dt, dw = (sg.DeviceArray(ctx, n) for _ in range(3))
    for j in range(kv - 1):
        dv.upload(np.ascontiguousarray(V[:, j]))
        A.spmv(dv, dt)
        M.apply(dt, dw)
        Y[:, j] = dw.get()
    # ||M^-1 A||_2 stands in for ||A||_2 here
    nrm = max(np.linalg.norm(Y[:, j]) for j in range(kv - 1))
    rel = _check_relation(Y, V, H, nrm)
    orth = _orth(V)
    print(f"gmres ras/fp64 {ortho}: ||V^T V - I|| = {orth:.2e}, relation {rel:.2e}")
    if ortho == "cgs2":
        assert orth <= 1e-12, orth


def test_basis_hook_errors(sg):
    c = sg.Context(0)
    with pytest.raises(sg.SgError):  # nothing solved on this context yet
        sg.last_basis(c)
    c.close()
