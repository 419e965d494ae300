"""Device convergence analytics (sg_estimate_condition, sg_estimate_spectral_radius)
against the reference's estimate_condition / estimate_spectral_radius
(src/analysis/condition_estimator.h, spectral_analysis.h; golden values from
oracle/_ref) and the CPU oracle's restatement (bit for bit against those
goldens in test_oracle.py).

Bars: Lanczos step counts equal, except where the recurrence reaches an
invariant subspace and the reference's beta at that step is itself roundoff
(below 1e-7 (1 + |alpha|), e.g. 2.6e-9 on Poisson 6 x 6): then the device may
stop there, earlier than the reference; Ritz extremes and kappa within 1e-9
relative (both recurrences are stable to about 1e-13 under a change of
summation order); power-iteration counts within +-1 and the radius within
1e-9 relative; the first Lanczos coefficients within 1e-12 of the oracle's.
"""
import numpy as np
import pytest

import oracle_bind as O
from conftest import golden

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sg():
    import __graft_entry__
    return __graft_entry__.load_package()


@pytest.fixture(scope="module")
def ctx(sg):
    c = sg.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("case", golden("analysis.json"), ids=lambda c: f"{c['pc']}-{c['grid']}")
def test_matches_reference(sg, ctx, case):
    A = O.grid(case["grid"])
    dA = sg.Csr(ctx, A.rp, A.ci, A.v)
    M = None if case["pc"] == "none" else sg.Precond(ctx, case["pc"]).setup(dA)
    Mo = None if case["pc"] == "none" else O.Precond(case["pc"], A)
    for key, MM, MMo in (("cond_A", None, None), ("cond_MA", M, Mo)):
        got, ref = sg.estimate_condition(dA, MM, case["k"]), case[key]
        if got["lanczos_steps"] != ref["steps"]:
            al, be = O.lanczos(A, MMo, case["k"])  # the reference's recurrence, bit for bit
            j = got["lanczos_steps"] - 1
            assert got["lanczos_steps"] < ref["steps"], key
            assert be[j] < 1e-7 * (1.0 + abs(al[j])), (key, j, be[j])
        assert got["valid"] == bool(ref["valid"]), key
        for f in ("lambda_min", "lambda_max", "kappa"):
            assert np.isclose(got[f], ref[f], rtol=1e-9, atol=0), (key, f, got[f], ref[f])
    s = sg.estimate_spectral_radius(dA, M, case["power_iters_max"])
    assert abs(s["power_iters"] - case["power_iters"]) <= 1
    assert np.isclose(s["spectral_radius"], case["rho"], rtol=1e-9, atol=0)
    assert s["convergence_ok"] == bool(case["ok"])


@pytest.mark.parametrize("pc,nparts", [("ras", 8), ("twolevel_ras", 8), ("asm", 4)])
def test_schwarz_preconditioners_match_oracle(sg, ctx, pc, nparts):
    A = O.grid("2d:48:48")
    dA = sg.Csr(ctx, A.rp, A.ci, A.v)
    M = sg.Precond(ctx, pc, nparts=nparts, overlap=1).setup(dA)
    Mo = O.Precond(pc, A, nparts, 1, M.partition())
    got = sg.estimate_condition(dA, M, 60)
    ref = O.condition(A, Mo, 60)
    assert got["lanczos_steps"] == ref["steps"]
    for f in ("lambda_min", "lambda_max"):
        assert np.isclose(got[f], ref[f], rtol=1e-9)
    al, be = O.lanczos(A, Mo, 60)
    assert np.allclose(got["alpha"][:8], al[:8], rtol=1e-12)
    assert np.allclose(got["beta"][:8], be[:8], rtol=1e-12)
    s = sg.estimate_spectral_radius(dA, M, 300)
    so = O.spectral(A, Mo, 300)
    assert abs(s["power_iters"] - so["power_iters"]) <= 1
    assert np.isclose(s["spectral_radius"], so["rho"], rtol=1e-9)


def test_closed_form_and_edge_cases(sg, ctx):
    # Poisson 2D 12 x 12: extreme eigenvalues 8 sin^2(pi/26) and 8 - that
    A = O.poisson2d(12, 12)
    dA = sg.Csr(ctx, A.rp, A.ci, A.v)
    c = sg.estimate_condition(dA, None, 80)
    lo = 8 * np.sin(np.pi / 26) ** 2
    assert np.isclose(c["lambda_min"], lo, rtol=1e-10)
    assert np.isclose(c["lambda_max"], 8 - lo, rtol=1e-10)
    assert np.isclose(c["kappa"], (8 - lo) / lo, rtol=1e-9) and c["valid"]
    # the tridiagonal's eigenvalues against LAPACK
    import scipy.linalg
    ev = scipy.linalg.eigh_tridiagonal(c["alpha"], c["beta"], eigvals_only=True)
    assert np.isclose(c["lambda_min"], ev[0], rtol=1e-12)
    assert np.isclose(c["lambda_max"], ev[-1], rtol=1e-12)
    # k < 2 or n < 2: no estimate
    assert sg.estimate_condition(dA, None, 1)["lanczos_steps"] == 0
    one = sg.Csr(ctx, np.array([0, 1], np.int32), np.array([0], np.int32), np.array([2.0]))
    assert sg.estimate_condition(one, None, 80)["lanczos_steps"] == 0
    # M = A^{-1} exactly (identity operator): E = 0 -> ||E v|| < 1e-14 at once
    n = 20
    I = sg.Csr(ctx, np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), np.ones(n))
    s = sg.estimate_spectral_radius(I, sg.Precond(ctx, "jacobi").setup(I), 50)
    assert s["power_iters"] == 1 and s["spectral_radius"] == 0.0 and s["convergence_ok"]
