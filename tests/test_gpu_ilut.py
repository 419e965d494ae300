"""Device ILUT (SG_PC_ILUT): the host factorisation restates the reference's
ILUTPrecond::setup (src/precond/ilut.h:28-149) and the device runs its
forward / backward substitution (ilut.h:152-172) as level-scheduled sweeps.

Bars: l_nnz / u_nnz equal and every FP64 apply bit for bit equal to the
reference's (golden vectors from oracle/_ref) and to the CPU oracle's on
larger matrices; GMRES / FGMRES iteration counts within +-1 of the
reference's; BiCGSTAB within its summation-order band; FP32 sweeps within
1e-5 of the oracle's FP32 restatement.
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


def dev(sg, ctx, a):
    return sg.DeviceArray.from_host(ctx, np.ascontiguousarray(a, np.float64))


def _apply(sg, ctx, M, x):
    dx, dy = dev(sg, ctx, x), dev(sg, ctx, np.zeros(len(x)))
    M.apply(dx, dy)
    return dy.get()


@pytest.mark.parametrize("case", golden("ilut.json")["applies"],
                         ids=lambda c: f"{c['grid']}-{c['drop_tol']}-{c['fill_factor']}")
def test_apply_bitwise_reference(sg, ctx, case):
    A = O.grid(case["grid"])
    dA = sg.Csr(ctx, A.rp, A.ci, A.v)
    M = sg.Precond(ctx, "ilut", drop_tol=case["drop_tol"],
                   fill_factor=case["fill_factor"]).setup(dA)
    assert M.ilut_nnz() == (case["l_nnz"], case["u_nnz"])
    assert np.array_equal(_apply(sg, ctx, M, np.array(case["x"])), np.array(case["y"]))


# 2d:512:512 has 1534 + 4088 sweep levels: too many for a level table in LDS,
# so its sweeps read the tables from global memory
@pytest.mark.parametrize("spec,drop,fill", [("2d:256:256", 1e-4, 10), ("3d:24:24:24", 1e-3, 4),
                                            ("rnd:20000:3", 1e-3, 8), ("2d:512:512", 1e-4, 10)])
def test_apply_bitwise_oracle(sg, ctx, spec, drop, fill):
    A = O.grid(spec)
    dA = sg.Csr(ctx, A.rp, A.ci, A.v)
    M = sg.Precond(ctx, "ilut", drop_tol=drop, fill_factor=fill).setup(dA)
    P = O.Precond("ilut", A, drop_tol=drop, fill_factor=fill)
    assert M.ilut_nnz() == (P.l_nnz, P.u_nnz)
    x = np.random.default_rng(3).standard_normal(A.n)
    assert np.array_equal(_apply(sg, ctx, M, x), P.apply(x))
    # FP32 sweeps against the oracle's FP32 restatement
    M32 = sg.Precond(ctx, "ilut", drop_tol=drop, fill_factor=fill, precision="fp32").setup(dA)
    P32 = O.Precond("ilut", A, drop_tol=drop, fill_factor=fill, precision=1)
    y32, r32 = _apply(sg, ctx, M32, x), P32.apply(x)
    assert np.linalg.norm(y32 - r32) <= 1e-5 * np.linalg.norm(r32)


@pytest.mark.parametrize("case", golden("ilut.json")["solves"], ids=lambda c: c["method"])
def test_solves_match_reference(sg, ctx, case):
    A = O.grid(case["grid"])
    if case["rhs"] == "ones":
        xs = np.ones(A.n)
    else:
        from rng_ref import splitmix_uniform
        xs = 2.0 * splitmix_uniform(int(case["rhs"].split(":")[1]), A.n) - 1.0
    b = O.spmv(A, xs)
    dA = sg.Csr(ctx, A.rp, A.ci, A.v)
    M = sg.Precond(ctx, "ilut").setup(dA)
    x = dev(sg, ctx, np.zeros(A.n))
    kw = dict(max_iter=case["max_iter"], tol=case["tol_arg"])
    if case["method"] != "bicgstab":
        kw["restart"] = case["restart"]
    r = sg.solve(dA, dev(sg, ctx, b), x, M, method=case["method"], **kw)
    band = 1 if case["method"] != "bicgstab" else max(2, round(0.12 * case["iterations"]))
    assert r["converged"] and abs(r["iterations"] - case["iterations"]) <= band
    assert np.linalg.norm(x.get() - xs) <= 1e-5 * np.linalg.norm(xs)


def test_defaults_and_errors(sg, ctx):
    A = O.grid("2d:16:16")
    dA = sg.Csr(ctx, A.rp, A.ci, A.v)
    M = sg.Precond(ctx, "ilut", drop_tol=0.0, fill_factor=0).setup(dA)  # -> 1e-4, 10
    P = O.Precond("ilut", A)
    assert M.ilut_nnz() == (P.l_nnz, P.u_nnz)
    with pytest.raises(sg.SgError):
        sg.Precond(ctx, "ilut", drop_tol=-1.0, fill_factor=10).setup(dA)
