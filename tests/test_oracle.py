"""Pin the CPU oracle before trusting it.

Each check is chosen so that a plausible restatement mistake (a dropped term,
a wrong sign or index, a transposed operand, a different summation order)
fails it: golden vectors from the reference compiled under oracle/_ref,
closed forms, library routines (scipy LU) and brute force on tiny inputs.
"""
import numpy as np
import pytest
import scipy.linalg
import scipy.sparse

import oracle_bind as O


def test_poisson_closed_forms():
    for nx, ny in ((1, 1), (3, 7), (16, 16)):
        A = O.poisson2d(nx, ny)
        n = nx * ny
        assert A.nnz == 5 * n - 2 * (nx + ny)
        # A*1: 4 minus the number of present neighbours (Dirichlet rows).
        ix, iy = np.arange(n) % nx, np.arange(n) // nx
        missing = (ix == 0).astype(int) + (ix == nx - 1) + (iy == 0) + (iy == ny - 1)
        assert np.array_equal(O.spmv(A, np.ones(n)), missing.astype(float))
        D = A.dense()
        assert np.array_equal(D, D.T)
    for nx, ny, nz in ((2, 3, 4), (6, 6, 6)):
        A = O.poisson3d(nx, ny, nz)
        n = nx * ny * nz
        assert A.nnz == 7 * n - 2 * (nx * ny + ny * nz + nx * nz)
        # columns strictly ascending per row, diagonal present
        for i in range(n):
            cols = A.ci[A.rp[i]:A.rp[i + 1]]
            assert np.all(np.diff(cols) > 0) and i in cols


def test_spmv_dense_and_sequential_order():
    A = O.poisson3d(5, 4, 3)
    rng = np.random.default_rng(0)
    x = rng.standard_normal(A.n)
    y = O.spmv(A, x)
    assert np.allclose(y, A.dense() @ x, rtol=0, atol=1e-13)
    # bitwise: CSR-order sequential sum starting from 0.0
    ref = np.array([sum((A.v[j] * x[A.ci[j]] for j in range(A.rp[i], A.rp[i + 1])), 0.0)
                    for i in range(A.n)])
    assert np.array_equal(y, ref)


def test_dot_closed_form():
    a = np.arange(1, 101, dtype=float)
    assert O.lib().or_dot(100, a, np.ones(100)) == 5050.0
    assert np.sqrt(O.lib().or_dot(100, np.ones(100), np.ones(100))) == 10.0


@pytest.mark.parametrize("spec", ["2d:8:8", "2d:16:16", "3d:6:6:6", "3d:5:4:3"])
def test_ilu0_matches_reference_bitwise(gold, spec):
    g = gold("ilu0.json")[spec]
    A = O.grid(spec)
    lu, diag = O.ilu0(A)
    assert np.array_equal(diag, np.array(g["diag"]))
    assert np.array_equal(lu, np.array(g["lu"]))
    y = O.ilu0_solve(A, lu, diag, np.array(g["x"]))
    assert np.array_equal(y, np.array(g["y"]))


def test_ilu0_pattern_identity_and_recurrence():
    A = O.poisson2d(9, 7)
    lu, diag = O.ilu0(A)
    n = A.n
    L = np.eye(n)
    U = np.zeros((n, n))
    for i in range(n):
        for j in range(A.rp[i], A.rp[i + 1]):
            c = A.ci[j]
            (L if c < i else U)[i, c] = lu[j]
    P = (A.dense() != 0)
    # (L U)_ij == A_ij on the pattern of A (defining property of ILU(0))
    assert np.allclose((L @ U)[P], A.dense()[P], atol=1e-13)
    # natural-order 5-point recurrence d_i = 4 - 1/d_{i-1} - 1/d_{i-nx}
    d = np.zeros(n)
    for i in range(n):
        t = 4.0
        if i % 9 > 0:
            t -= 1.0 / d[i - 1]
        if i >= 9:
            t -= 1.0 / d[i - 9]
        d[i] = t
    assert np.allclose(lu[diag], d, rtol=1e-14)


def test_ilu0_exact_on_tridiagonal():
    A = O.poisson2d(40, 1)  # tridiagonal: ILU(0) is the exact LU
    lu, diag = O.ilu0(A)
    b = np.random.default_rng(1).standard_normal(A.n)
    x = O.ilu0_solve(A, lu, diag, b)
    assert np.allclose(x, np.linalg.solve(A.dense(), b), atol=1e-12)


def test_overlap_sets_match_bfs_fixtures(gold):
    for case in gold("blocks.json"):
        A = O.grid(case["grid"])
        sets = O.overlap_sets(A, case["part"], case["nparts"], case["overlap"])
        for s, ref in zip(sets, case["sets"]):
            assert np.array_equal(s, np.array(ref))


def test_ras_asm_match_reference_blocks(gold):
    for case in gold("blocks.json"):
        A = O.grid(case["grid"])
        x = np.array(case["x"])
        for kind in ("ras", "asm"):
            pc = O.Precond(kind, A, case["nparts"], case["overlap"], case["part"])
            y = pc.apply(x)
            assert np.array_equal(y, np.array(case[kind])), (case["grid"], kind)


def test_ras_special_cases():
    A = O.poisson2d(12, 10)
    x = np.random.default_rng(2).standard_normal(A.n)
    # one subdomain == global ILU(0)
    y1 = O.Precond("ras", A, 1, 2).apply(x)
    assert np.array_equal(y1, O.Precond("ilu0", A).apply(x))
    # overlap 0: RAS == ASM == block ILU(0)
    part = (np.arange(A.n) % 12 >= 6).astype(np.int32) + 2 * (np.arange(A.n) // 12 >= 5)
    yr = O.Precond("ras", A, 4, 0, part).apply(x)
    ya = O.Precond("asm", A, 4, 0, part).apply(x)
    assert np.array_equal(yr, ya)
    D = A.dense()
    yb = np.zeros(A.n)
    for s in range(4):
        idx = np.flatnonzero(part == s)
        sub = O.Csr(*_extract(A, idx))
        lu, dg = O.ilu0(sub)
        yb[idx] = O.ilu0_solve(sub, lu, dg, x[idx])
    assert np.array_equal(yr, yb)
    del D


def _extract(A, idx):
    inv = -np.ones(A.n, int)
    inv[idx] = np.arange(len(idx))
    rp, ci, v = [0], [], []
    for gi in idx:
        for j in range(A.rp[gi], A.rp[gi + 1]):
            if inv[A.ci[j]] >= 0:
                ci.append(inv[A.ci[j]])
                v.append(A.v[j])
        rp.append(len(ci))
    return rp, ci, v


def _dense_two_level(A, part, nparts):
    D = A.dense()
    R0 = np.zeros((nparts, A.n))
    for s in range(nparts):
        own = part == s
        R0[s, own] = 1.0 / np.sqrt(own.sum())
    return R0, R0 @ D @ R0.T


@pytest.mark.parametrize("fine", ["twolevel_ras", "twolevel_asm"])
def test_two_level_matches_dense_form(fine):
    A = O.poisson2d(14, 13)
    rng = np.random.default_rng(3)
    part = ((np.arange(A.n) % 14) // 5 + 3 * ((np.arange(A.n) // 14) // 5)).astype(np.int32)
    nparts = int(part.max()) + 1
    pc = O.Precond(fine, A, nparts, 1, part)
    A0, LU, piv = pc.coarse()
    R0, A0_dense = _dense_two_level(A, part, nparts)
    assert np.allclose(A0, A0_dense, rtol=0, atol=1e-13)
    # Galerkin coarse matrix of an SPD operator is SPD
    assert np.allclose(A0, A0.T, atol=1e-13)
    x = rng.standard_normal(A.n)
    fine_y = O.Precond(fine[len("twolevel_"):], A, nparts, 1, part).apply(x)
    want = fine_y + R0.T @ np.linalg.solve(A0_dense, R0 @ x)
    assert np.allclose(pc.apply(x), want, rtol=0, atol=1e-12)


def test_lu_matches_reference_and_scipy(gold):
    g = gold("lu.json")
    for key in ("n16", "n40"):
        n = int(round(np.sqrt(len(g[key]["A"]))))
        A = np.array(g[key]["A"]).reshape(n, n)
        rc, LU, piv = O.lu_factor(A)
        assert rc == 0
        x = O.lu_solve(LU, piv, np.array(g[key]["b"]))
        assert np.array_equal(x, np.array(g[key]["x"]))
        # same pivot choice as LAPACK getrf (first max |.|), values to rounding
        lu_s, piv_s = scipy.linalg.lu_factor(A)
        assert np.array_equal(piv, piv_s)
        assert np.allclose(LU, lu_s, rtol=1e-12, atol=1e-13)
    n = 8
    A = np.array(g["singular"].get("A", [0.0] * 64)).reshape(n, n) if "A" in g["singular"] else None
    assert g["singular"] == {"error": "Coarse matrix is singular"}
    S = np.zeros((4, 4))
    rc, _, _ = O.lu_factor(S)
    assert rc == -1
    del A


@pytest.mark.parametrize("idx", range(12))
def test_krylov_matches_reference_golden(gold, idx):
    g = gold("solve.json")[idx]
    A = O.grid(g["grid"])
    n = A.n
    if g["rhs"] == "ones":
        xs = np.ones(n)
    else:
        xs = _uniform(int(g["rhs"].split(":")[1]), n)
    b = O.spmv(A, xs)
    pc = None if g["pc"] == "none" else O.Precond(g["pc"], A)
    r = O.solve(g["method"], A, b, pc, g["restart"], g["max_iter"], g["tol"])
    assert r["iterations"] == g["iterations"]
    assert r["converged"] == bool(g["converged"])
    assert np.array_equal(r["hist"], np.array(g["hist"]))
    assert r["residual_norm"] == g["residual_norm"]
    if "x" in g:
        assert np.array_equal(r["x"], np.array(g["x"]))


def _uniform(seed, n):
    from rng_ref import splitmix_uniform  # noqa: E402
    return 2.0 * splitmix_uniform(seed, n) - 1.0


def test_known_answers_exactness():
    # identity: lucky breakdown after one step, x = b (tests/test_gmres.cpp:27-44 spec)
    n = 10
    I = O.Csr(np.arange(n + 1), np.arange(n), np.ones(n))
    b = np.arange(1, n + 1, dtype=float)
    for m in ("gmres", "fgmres"):
        r = O.solve(m, I, b)
        assert r["converged"] and r["iterations"] <= 2
        assert np.allclose(r["x"], b, atol=1e-12)
    # 3x3 SPD with x* = [1,2,3] under none/jacobi/ilu0 (tests/test_gmres.cpp:46-92 spec)
    S = scipy.sparse.csr_matrix(np.array([[4.0, 1, 0], [1, 3, 1], [0, 1, 2]]))
    A = O.Csr(S.indptr, S.indices, S.data)
    xs = np.array([1.0, 2, 3])
    b = O.spmv(A, xs)
    for kind in ("none", "jacobi", "ilu0"):
        pc = None if kind == "none" else O.Precond(kind, A)
        r = O.solve("gmres", A, b, pc)
        assert np.abs(r["x"] - xs).sum() < 1e-8


def test_fgmres_estimate_is_true_residual():
    A = O.poisson2d(24, 24)
    b = O.spmv(A, _uniform(9, A.n))
    pc = O.Precond("ilu0", A)
    r = O.solve("fgmres", A, b, pc, 30, 1000, 1e-9)
    true = np.linalg.norm(b - O.spmv(A, r["x"]))
    assert abs(true - r["residual_norm"]) <= 1e-6 * max(true, 1e-12) + 1e-12


def test_metrics_json_golden(gold):
    lines = gold("metrics.json")
    assert O.metrics_json(matrix="poisson2d_256", method="GMRES+ilu0", n=65536, nnz=326656,
                          threads=8, overlap=1, iterations=412, tol=1e-8,
                          residual_norm=9.87654321e-9, time_setup_s=0.0123456789,
                          time_solve_s=1.5) == lines[0]
    assert O.metrics_json() == lines[1]
    assert O.metrics_json(matrix="poisson3d_128", method="GPU_GMRES+ras", n=2097152,
                          nnz=14581760, gpus=1, restart=50, tol=0.000123, residual_norm=0.0,
                          time_solve_s=123456.5, speedup=3.25, efficiency=1e-7) == lines[2]


def test_mixed_precision_oracle_close_to_fp64():
    A = O.poisson2d(20, 20)
    part = ((np.arange(A.n) % 20) // 10 + 2 * ((np.arange(A.n) // 20) // 10)).astype(np.int32)
    x = np.random.default_rng(4).standard_normal(A.n)
    for kind in ("ras", "twolevel_ras"):
        y64 = O.Precond(kind, A, 4, 1, part, 0).apply(x)
        y32 = O.Precond(kind, A, 4, 1, part, 1).apply(x)
        assert np.linalg.norm(y32 - y64) <= 1e-5 * np.linalg.norm(y64)
        assert not np.array_equal(y32, y64)


# ---- BiCGSTAB and the analysis routines -------------------------------------------

def _rhs(A, rhs):
    xs = np.ones(A.n) if rhs == "ones" else _uniform(int(rhs.split(":")[1]), A.n)
    return O.spmv(A, xs)


@pytest.mark.parametrize("case", __import__("conftest").golden("bicgstab.json"),
                         ids=lambda c: f"{c['pc']}-{c['grid']}-{c['max_iter']}")
def test_bicgstab_bitwise_reference(case):
    """or_bicgstab reproduces the reference bicgstab (bicgstab.h) bit for bit."""
    A = O.grid(case["grid"])
    b = _rhs(A, case["rhs"])
    pc = None if case["pc"] == "none" else O.Precond(case["pc"], A)
    r = O.bicgstab(A, b, pc, case["max_iter"], case["tol_arg"])
    assert r["iterations"] == case["iterations"]
    assert r["converged"] == bool(case["converged"])
    assert r["residual_norm"] == case["residual_norm"]
    assert np.array_equal(r["hist"], np.array(case["hist"]))
    if "x" in case:
        assert np.array_equal(r["x"], np.array(case["x"]))


def test_bicgstab_exact_cases():
    # identity: one half step (s = 0), x = b
    n = 7
    I = O.Csr(np.arange(n + 1), np.arange(n), np.ones(n))
    b = np.arange(1.0, n + 1)
    r = O.bicgstab(I, b)
    assert (r["iterations"], r["converged"], r["residual_norm"]) == (1, True, 0.0)
    assert np.array_equal(r["x"], b) and r["hist"].tolist() == [0.0]
    # 3x3 SPD with x* = [1, 2, 3] (the reference's GMRES test system): BiCGSTAB in
    # exact arithmetic terminates in at most 3 iterations; in floating point it
    # reaches the tolerance quickly and solves it
    A = O.Csr([0, 2, 5, 7], [0, 1, 0, 1, 2, 1, 2], [4.0, 1.0, 1.0, 3.0, 1.0, 1.0, 2.0])
    xs = np.array([1.0, 2.0, 3.0])
    r = O.bicgstab(A, O.spmv(A, xs), tol=1e-12)
    assert r["converged"] and r["iterations"] <= 4
    assert np.allclose(r["x"], xs, rtol=0, atol=1e-10)


def test_tridiag_eigs_against_lapack():
    rng = np.random.default_rng(4)
    for m in (2, 3, 10, 80):
        d, e = rng.standard_normal(m), rng.standard_normal(m - 1)
        ref = scipy.linalg.eigh_tridiagonal(d, e, eigvals_only=True)
        assert np.allclose(O.tridiag_eigs(d, e), ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    # 1D Laplacian: 2 - 2 cos(k pi / (m + 1))
    m = 50
    got = O.tridiag_eigs(np.full(m, 2.0), np.full(m - 1, -1.0))
    exact = 2.0 - 2.0 * np.cos(np.arange(1, m + 1) * np.pi / (m + 1))
    assert np.allclose(got, exact, rtol=0, atol=1e-13)


@pytest.mark.parametrize("case", __import__("conftest").golden("analysis.json"),
                         ids=lambda c: f"{c['pc']}-{c['grid']}")
def test_condition_and_spectral_bitwise_reference(case):
    A = O.grid(case["grid"])
    pc = None if case["pc"] == "none" else O.Precond(case["pc"], A)
    for key, M in (("cond_A", None), ("cond_MA", pc)):
        got, ref = O.condition(A, M, case["k"]), case[key]
        for f in ("lambda_min", "lambda_max", "kappa", "steps", "valid"):
            assert got[f] == ref[f], (key, f)
    s = O.spectral(A, pc, case["power_iters_max"])
    assert (s["rho"], s["power_iters"], s["ok"]) == (case["rho"], case["power_iters"], case["ok"])


def test_lanczos_closed_form_spectrum():
    # Poisson 2D on 12 x 12: 80 steps resolve the extreme eigenvalues
    # 4 - 2cos(pi i/13) - 2cos(pi j/13)
    A = O.poisson2d(12, 12)
    c = O.condition(A, None, 80)
    lo = 8 * np.sin(np.pi / 26) ** 2
    hi = 8 - lo
    assert np.isclose(c["lambda_min"], lo, rtol=1e-10) and np.isclose(c["lambda_max"], hi,
                                                                       rtol=1e-10)
    al, be = O.lanczos(A, None, 5)
    # alpha_0 = v0' A v0 for the LCG start vector
    from rng_ref import lcg_unit
    v0 = lcg_unit(A.n, 42)
    assert np.isclose(al[0], v0 @ A.dense() @ v0, rtol=1e-13)


def _bicgstab_numpy(A, b, pc, max_iter, tol, dot):
    """bicgstab.h's steps with a caller-chosen inner product (summation order)."""
    x = np.zeros(A.n)
    r = b - O.spmv(A, x)
    rh = r.copy()
    r0 = np.sqrt(dot(r, r))
    rho = al = om = 1.0
    p, v = np.zeros(A.n), np.zeros(A.n)
    for it in range(1, max_iter + 1):
        rp, rho = rho, dot(rh, r)
        p = r.copy() if it == 1 else r + ((rho / rp) * (al / om)) * (p - om * v)
        y = pc.apply(p)
        v = O.spmv(A, y)
        al = rho / dot(rh, v)
        s = r - al * v
        if np.sqrt(dot(s, s)) < tol * r0:
            return it
        z = pc.apply(s)
        t = O.spmv(A, z)
        om = dot(t, s) / dot(t, t)
        x = x + (al * y + om * z)
        r = s - om * t
        if np.sqrt(dot(r, r)) < tol * r0:
            return it
    return max_iter


def test_bicgstab_count_depends_on_summation_order():
    """Why the device BiCGSTAB is held to a band, not +-1: on Poisson 100^2 with
    ILU(0) the reference's 67 iterations become 74 with numpy's pairwise dot
    and 71 with extended-precision sums -- the same algorithm, other rounding."""
    A = O.poisson2d(100, 100)
    b = O.spmv(A, np.ones(A.n))
    pc = O.Precond("ilu0", A)
    seq = O.bicgstab(A, b, pc, 1000, 1e-10)["iterations"]
    pair = _bicgstab_numpy(A, b, pc, 1000, 1e-10, np.dot)
    ext = _bicgstab_numpy(A, b, pc, 1000, 1e-10,
                          lambda a, c: float(np.sum(a.astype(np.longdouble) * c)))
    assert seq == 67 and pair != seq and ext != seq
    band = max(2, round(0.12 * seq))
    assert abs(pair - seq) <= band and abs(ext - seq) <= band


# ---- ILUT ----------------------------------------------------------------------------

@pytest.mark.parametrize("case", __import__("conftest").golden("ilut.json")["applies"],
                         ids=lambda c: f"{c['grid']}-{c['drop_tol']}-{c['fill_factor']}")
def test_ilut_apply_bitwise_reference(case):
    A = O.grid(case["grid"])
    P = O.Precond("ilut", A, drop_tol=case["drop_tol"], fill_factor=case["fill_factor"])
    assert (P.l_nnz, P.u_nnz) == (case["l_nnz"], case["u_nnz"])
    assert np.array_equal(P.apply(np.array(case["x"])), np.array(case["y"]))


@pytest.mark.parametrize("case", __import__("conftest").golden("ilut.json")["solves"],
                         ids=lambda c: c["method"])
def test_ilut_solves_bitwise_reference(case):
    A = O.grid(case["grid"])
    b = _rhs(A, case["rhs"])
    P = O.Precond("ilut", A)
    if case["method"] == "bicgstab":
        r = O.bicgstab(A, b, P, case["max_iter"], case["tol_arg"])
    else:
        r = O.solve(case["method"], A, b, P, case["restart"], case["max_iter"], case["tol_arg"])
    assert r["iterations"] == case["iterations"]
    assert np.array_equal(r["hist"], np.array(case["hist"]))


def test_ilut_exact_on_tridiagonal():
    """No dropping and enough fill: ILUT of a tridiagonal matrix is its exact LU."""
    n = 40
    rng = np.random.default_rng(8)
    d = 4 + rng.random(n)
    e = rng.standard_normal(n - 1)
    f = rng.standard_normal(n - 1)
    T = np.diag(d) + np.diag(e, -1) + np.diag(f, 1)
    rows, cols = np.nonzero(T)
    rp = np.searchsorted(rows, np.arange(n + 1))
    A = O.Csr(rp, cols, T[rows, cols])
    P = O.Precond("ilut", A, drop_tol=0.0, fill_factor=5)
    x = rng.standard_normal(n)
    assert np.allclose(P.apply(x), np.linalg.solve(T, x), rtol=1e-12, atol=1e-12)
    assert P.l_nnz == n - 1 and P.u_nnz == 2 * n - 1
