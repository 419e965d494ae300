"""Parity of the HIP path (through the C ABI) with the CPU oracle and the
reference's golden vectors.

Bars: bit-exact for SpMV, ILU(0) factors and the FP64 sweeps (same operation
order as the reference, no FMA contraction); iteration counts within +-1 and
solutions within 1e-6 relative for GMRES/FGMRES (reductions are trees on the
device, sequential sums on the CPU); 1e-12 relative for FP64 preconditioner
applies that contain reductions (coarse restriction); 1e-5 for the FP32 modes.
"""
import numpy as np
import pytest

import oracle_bind as O

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


def test_library_exports_every_declared_symbol(sg):
    import re
    import os
    hdr = open(os.path.join(os.path.dirname(__file__), "..", "include", "sg_capi.h")).read()
    declared = set(re.findall(r"\b(sg_[a-z0-9_]+)\s*\(", hdr))
    for name in declared:
        assert hasattr(sg.lib(), name), name


@pytest.mark.parametrize("spec", ["2d:33:17", "3d:12:9:7", "2d:256:256"])
def test_spmv_bitwise(sg, ctx, spec):
    A = O.grid(spec)
    rng = np.random.default_rng(1)
    v = A.v * rng.uniform(0.5, 1.5, A.nnz)  # non-trivial values
    Ao = O.Csr(A.rp, A.ci, v)
    x = rng.standard_normal(A.n)
    dA = sg.Csr(ctx, A.rp, A.ci, v)
    dx, dy = dev(sg, ctx, x), dev(sg, ctx, np.zeros(A.n))
    dA.spmv(dx, dy)
    assert np.array_equal(dy.get(), O.spmv(Ao, x))
    # r = b - A x form used at every restart (alpha = -1, beta = 1)
    b = rng.standard_normal(A.n)
    db = dev(sg, ctx, b)
    dA.spmv(dx, db, alpha=-1.0, beta=1.0)
    assert np.array_equal(db.get(), b - O.spmv(Ao, x))


def _spmv_cases():
    rng = np.random.default_rng(7)
    out = []
    A = O.grid("3d:40:36:30")
    out.append(("3d", A.rp, A.ci, A.v * rng.uniform(0.5, 1.5, A.nnz), None))
    A = O.grid("rnd:3000:5")
    out.append(("rnd", A.rp, A.ci, A.v, None))
    # ragged rows (0..12 entries, unsorted lengths) and a batch of systems
    # whose sizes leave partial 64-row slices at every system end
    sizes = [70, 1, 130, 64, 257]
    rows, off = [], 0
    for m in sizes:
        for i in range(m):
            k = int(rng.integers(0, min(m, 12) + 1))
            rows.append(sorted(off + rng.choice(m, size=k, replace=False)))
        off += m
    rp = np.cumsum([0] + [len(r) for r in rows]).astype(np.int32)
    ci = np.concatenate([np.array(r, np.int32) for r in rows])
    out.append(("ragged_batch", rp, ci, rng.standard_normal(len(ci)), np.cumsum([0] + sizes)))
    # couplings 40000 rows away: no 16-bit offset copy, mode 6 runs CSR
    n = 70000
    rows = [sorted({i, (i + 40000) % n, (i + 1) % n}) for i in range(n)]
    rp = np.cumsum([0] + [len(r) for r in rows]).astype(np.int32)
    ci = np.concatenate([np.array(r, np.int32) for r in rows])
    out.append(("far", rp, ci, rng.standard_normal(len(ci)), None))
    return out


@pytest.mark.parametrize("case", _spmv_cases(), ids=lambda c: c[0])
def test_spmv_every_shape_bitwise(sg, case):
    """Each SpMV shape (CSR with 256 / 512 / 1024 rows per workgroup and the
    SELL-64 copy with 32-bit columns or 16-bit offsets, forced through autotune
    modes 2..6) equals the serial row loop bit for bit, signed zeros included,
    in the y = A x and y = b - A x forms."""
    _, rp, ci, v, sys_rows = case
    n = len(rp) - 1
    rng = np.random.default_rng(8)
    x = rng.standard_normal(n)
    x[::7] = -0.0
    b = rng.standard_normal(n)
    Ao = O.Csr(rp, ci, v)
    ref1, ref2 = O.spmv(Ao, x), b - O.spmv(Ao, x)
    for mode in (2, 3, 4, 5, 6):
        c = sg.Context(0)
        c.autotune(mode)
        if sys_rows is None:
            dA = sg.Csr(c, rp, ci, v)
        else:
            mats = []
            for q in range(len(sys_rows) - 1):
                a, e = sys_rows[q], sys_rows[q + 1]
                lrp = rp[a:e + 1] - rp[a]
                mats.append((lrp, ci[rp[a]:rp[e]] - a, v[rp[a]:rp[e]]))
            dA = sg.BatchCsr(c, mats)
        dx, dy = dev(sg, c, x), dev(sg, c, np.zeros(n))
        dA.spmv(dx, dy)
        assert np.array_equal(dy.get().view(np.int64), ref1.view(np.int64)), mode
        db = dev(sg, c, b)
        dA.spmv(dx, db, alpha=-1.0, beta=1.0)
        assert np.array_equal(db.get().view(np.int64), ref2.view(np.int64)), mode
        del dA, dx, dy, db
        c.close()


def test_spmv_long_rows(sg):
    """Rows longer than a workgroup's LDS staging capacity (2048 / 4096 / 8192
    products at 256 / 512 / 1024 threads): 3000 and 10000 entries, next to
    empty and one-entry rows.  Every CSR shape (autotune modes 2..6; the SELL
    copy is not built for such rows) sums them in CSR order, bit for bit equal
    to the serial row loop (src/core/sparse_matrix.h:30-37)."""
    n = 12000
    rng = np.random.default_rng(2)
    rows = [[0, 1], list(range(3000)), [], [3], sorted(rng.choice(n, 10000, replace=False))]
    rows += [[i] for i in range(len(rows), n)]
    rp = np.cumsum([0] + [len(r) for r in rows]).astype(np.int32)
    ci = np.concatenate([np.array(r, np.int32) for r in rows if r])
    v = rng.standard_normal(len(ci)) * np.exp(rng.uniform(-20, 20, len(ci)))
    x = np.random.default_rng(3).standard_normal(n)
    ref = O.spmv(O.Csr(rp, ci, v), x)
    for mode in (2, 3, 4, 5, 6):
        c = sg.Context(0)
        c.autotune(mode)
        dA = sg.Csr(c, rp, ci, v)
        dx, dy = dev(sg, c, x), dev(sg, c, np.zeros(n))
        dA.spmv(dx, dy)
        assert np.array_equal(dy.get().view(np.int64), ref.view(np.int64)), mode
        del dA, dx, dy
        c.close()


def test_blas1_closed_forms(sg, ctx):
    a = dev(sg, ctx, np.arange(1, 101, dtype=float))
    ones = dev(sg, ctx, np.ones(100))
    import ctypes as C
    r = C.c_double()
    ctx.check(sg.lib().sg_dot(ctx.h, 100, a.ptr, ones.ptr, C.byref(r)))
    assert r.value == 5050.0
    ctx.check(sg.lib().sg_nrm2(ctx.h, 100, ones.ptr, C.byref(r)))
    assert r.value == 10.0
    x = np.random.default_rng(4).standard_normal(3_000_001)
    dx = dev(sg, ctx, x)
    ctx.check(sg.lib().sg_nrm2(ctx.h, x.size, dx.ptr, C.byref(r)))
    assert abs(r.value - np.linalg.norm(x)) <= 1e-13 * np.linalg.norm(x)


@pytest.mark.parametrize("spec", ["2d:8:8", "2d:16:16", "3d:6:6:6", "3d:5:4:3"])
def test_ilu0_factors_bitwise_golden(sg, ctx, gold, spec):
    g = gold("ilu0.json")[spec]
    A = O.grid(spec)
    M = sg.Precond(ctx, "ilu0").setup(sg.Csr(ctx, A.rp, A.ci, A.v))
    assert np.array_equal(M.ilu_values(0), np.array(g["lu"]))
    x, y = dev(sg, ctx, g["x"]), dev(sg, ctx, np.zeros(A.n))
    M.apply(x, y)
    assert np.array_equal(y.get(), np.array(g["y"]))


def test_ilu0_random_values_bitwise(sg, ctx):
    A = O.poisson3d(20, 18, 16)
    rng = np.random.default_rng(5)
    v = A.v * rng.uniform(0.8, 1.2, A.nnz)
    Ao = O.Csr(A.rp, A.ci, v)
    lu, diag = O.ilu0(Ao)
    M = sg.Precond(ctx, "ilu0").setup(sg.Csr(ctx, A.rp, A.ci, v))
    assert np.array_equal(M.ilu_values(0), lu)
    x = rng.standard_normal(A.n)
    dx, dy = dev(sg, ctx, x), dev(sg, ctx, np.zeros(A.n))
    M.apply(dx, dy)
    assert np.array_equal(dy.get(), O.ilu0_solve(Ao, lu, diag, x))


@pytest.mark.parametrize("spec,kind,nparts", [
    ("2d:256:256", "ilu0", 1),     # 65536 FP64 rows: window of recent values in LDS
    ("rnd:60000:3", "ilu0", 1),    # random couplings: many references outside the window
    ("3d:48:48:48", "ras", 2),     # two subdomains of ~57K rows each
    ("3d:40:40:40", "asm", 2),
])
def test_ring_sweep_bitwise(sg, ctx, spec, kind, nparts):
    """Subdomains whose solution vector exceeds LDS run the ring sweep, or its
    staged form (records and own values streamed through LDS, modes 9..11) when
    every level has at most four entries per row and every reference stays in
    the ring: both must equal the oracle bit for bit in FP64 with every
    workgroup size."""
    A = O.grid(spec)
    rng = np.random.default_rng(12)
    v = A.v * rng.uniform(0.9, 1.1, A.nnz)
    Ao = O.Csr(A.rp, A.ci, v)
    dA = sg.Csr(ctx, A.rp, A.ci, v)
    x = rng.standard_normal(A.n)
    outs = []
    staged = []
    for mode in (2, 3, 4, 7, 8, 9, 10, 11):  # 256 / 512 / 1024 threads; deep prefetch; staged
        ctx.autotune(mode)
        M = sg.Precond(ctx, kind, nparts=nparts, overlap=1).setup(dA)
        info = M.stats()
        assert not info["z_in_lds"]
        y0 = dev(sg, ctx, np.zeros(A.n))
        M.apply(dev(sg, ctx, x), y0)
        staged.append(M.stats()["sweep_shape"] == mode - 2)
        y = dev(sg, ctx, np.zeros(A.n))
        M.apply(dev(sg, ctx, x), y)
        outs.append(y.get())
        part = M.partition() if nparts > 1 else None
    ctx.autotune(1)
    ref = O.Precond(kind, Ao, nparts, 1, part).apply(x)
    for y in outs:
        assert np.array_equal(y, ref)
    # the stencil cases stream through the staged ring sweep (the random one
    # has rows with more than four entries per level and falls back)
    assert all(staged[5:]) == (spec != "rnd:60000:3"), staged


def _stencil9(nx, ny):
    """2D 9-point stencil (four lower and four upper neighbours per row)."""
    rows = []
    for j in range(ny):
        for i in range(nx):
            r = []
            for dj in (-1, 0, 1):
                for di in (-1, 0, 1):
                    a, b = i + di, j + dj
                    if 0 <= a < nx and 0 <= b < ny:
                        r.append((b * nx + a, 8.0 if (di, dj) == (0, 0) else -1.0))
            rows.append(r)
    rp = np.cumsum([0] + [len(r) for r in rows]).astype(np.int32)
    ci = np.array([c for r in rows for c, _ in r], np.int32)
    v = np.array([x for r in rows for _, x in r])
    return O.Csr(rp, ci, v)


@pytest.mark.parametrize("spec,kind,nparts,prec", [
    ("2d:100:100", "ilu0", 1, "fp64"),   # 199 levels per sweep in ~18 LDS chunks
    ("3d:24:24:24", "ras", 8, "fp64"),
    ("3d:24:24:24", "asm", 8, "fp64"),
    ("3d:30:30:30", "ilu0", 1, "fp32"),  # levels of up to 675 rows (> 512 threads)
    # z takes 108 KiB of LDS, leaving two 24 KiB buffers of 384 FP64 records:
    # the widest levels (432 rows) are split across chunks
    ("3d:24:24:24", "ilu0", 1, "fp64"),
    # FP32 with at most three entries per row (the stencils above): 24-byte
    # records; a 9-point stencil has four, so it keeps the 32-byte records
    ("9pt:90:90", "ilu0", 1, "fp32"),
    ("9pt:90:90", "ras", 4, "fp32"),
    # FP32 with at most two entries per row (5-point): one-word forward records
    ("2d:100:100", "ilu0", 1, "fp32"),
    ("2d:120:120", "ras", 4, "fp32"),
])
def test_staged_sweep_bitwise(sg, ctx, spec, kind, nparts, prec):
    """The staged sweeps (factor data copied into LDS a chunk of levels ahead;
    autotune modes 9, 10 and 11) equal the barrier sweep bit for bit, and the
    oracle bit for bit in FP64."""
    A = _stencil9(*(int(t) for t in spec.split(":")[1:])) if spec.startswith("9pt") else O.grid(spec)
    rng = np.random.default_rng(14)
    v = A.v * rng.uniform(0.9, 1.1, A.nnz)
    Ao = O.Csr(A.rp, A.ci, v)
    dA = sg.Csr(ctx, A.rp, A.ci, v)
    x = rng.standard_normal(A.n)
    outs = {}
    for mode in (3, 9, 10, 11):
        ctx.autotune(mode)
        M = sg.Precond(ctx, kind, nparts=nparts, overlap=1, precision=prec).setup(dA)
        y = dev(sg, ctx, np.zeros(A.n))
        M.apply(dev(sg, ctx, x), y)
        info = M.stats()
        assert info["z_in_lds"] and info["sweep_shape"] == mode - 2, (mode, info)
        outs[mode] = y.get()
        part = M.partition() if nparts > 1 else None
    ctx.autotune(1)
    for mode in (9, 10, 11):
        assert np.array_equal(outs[mode], outs[3]), mode
    if prec == "fp64":
        assert np.array_equal(outs[9], O.Precond(kind, Ao, nparts, 1, part).apply(x))


def test_sweep_clock(sg, ctx):
    """The staged sweep's workgroup clock (sg_precond_sweep_clock): the clocked
    apply gives the plain apply's values bit for bit, and every subdomain's
    workgroup records once, with its own size; a sweep that is not the staged
    kernel is refused."""
    A = O.grid("2d:120:120")
    dA = sg.Csr(ctx, A.rp, A.ci, A.v)
    x = dev(sg, ctx, np.random.default_rng(15).standard_normal(A.n))
    ctx.autotune(10)
    M = sg.Precond(ctx, "twolevel_ras", nparts=16, overlap=1, precision="fp32",
                   coarse_solve=1).setup(dA)
    y, yc = dev(sg, ctx, np.zeros(A.n)), dev(sg, ctx, np.zeros(A.n))
    M.apply(x, y)
    assert M.stats()["sweep_shape"] == 8
    r = M.sweep_clock(x, yc)
    assert np.array_equal(yc.get(), y.get())
    assert sorted(r["sub"].tolist()) == list(range(16))
    sizes = [len(d) for d in M.subdomains()]
    assert [int(r["rows"][k]) for k in np.argsort(r["sub"])] == sizes
    assert r["start_us"].min() == 0.0
    assert np.all(r["start_us"] <= r["gathered_us"]) and np.all(r["gathered_us"] < r["swept_us"])
    assert np.all(r["swept_us"] <= r["end_us"])
    assert r["owned"].sum() == A.n and np.all(r["owned"] <= r["rows"])
    assert np.all(r["flev"] > 0) and np.all(r["blev"] > 0) and np.all(r["chunks"] > 0)
    assert r["rate_khz"] > 0
    ctx.autotune(3)
    Mb = sg.Precond(ctx, "ras", nparts=16, overlap=1, precision="fp32",
                    partition=M.partition()).setup(dA)
    with pytest.raises(sg.SgError):
        Mb.sweep_clock(x, yc)
    ctx.autotune(1)


def test_ring_sweep_fp32(sg, ctx):
    A = O.poisson3d(64, 64, 64)
    dA = sg.Csr(ctx, A.rp, A.ci, A.v)
    M = sg.Precond(ctx, "ras", nparts=4, overlap=1, precision="fp32").setup(dA)
    assert not M.stats()["z_in_lds"]
    x = np.random.default_rng(13).standard_normal(A.n)
    y = dev(sg, ctx, np.zeros(A.n))
    M.apply(dev(sg, ctx, x), y)
    ref = O.Precond("ras", A, 4, 1, M.partition(), precision=1).apply(x)
    assert np.linalg.norm(y.get() - ref) <= 1e-5 * np.linalg.norm(ref)
    y2 = dev(sg, ctx, np.zeros(A.n))
    M.apply(dev(sg, ctx, x), y2)
    assert np.array_equal(y.get(), y2.get())
    # the staged ring sweep gives the same values bit for bit
    for mode in (9, 10, 11):
        ctx.autotune(mode)
        Ms = sg.Precond(ctx, "ras", nparts=4, overlap=1, precision="fp32",
                        partition=M.partition()).setup(dA)
        ys = dev(sg, ctx, np.zeros(A.n))
        Ms.apply(dev(sg, ctx, x), ys)
        assert Ms.stats()["sweep_shape"] == mode - 2
        assert np.array_equal(ys.get(), y.get()), mode
    ctx.autotune(1)


def test_ras_asm_match_reference_blocks(sg, ctx, gold):
    for case in gold("blocks.json"):
        A = O.grid(case["grid"])
        dA = sg.Csr(ctx, A.rp, A.ci, A.v)
        x = dev(sg, ctx, case["x"])
        for kind in ("ras", "asm"):
            M = sg.Precond(ctx, kind, case["nparts"], case["overlap"],
                           partition=case["part"]).setup(dA)
            for s, ref in zip(M.subdomains(), case["sets"]):
                assert np.array_equal(s, np.array(ref))
            y = dev(sg, ctx, np.zeros(A.n))
            M.apply(x, y)
            assert np.array_equal(y.get(), np.array(case[kind])), (case["grid"], kind)


def test_partitioner_properties_and_overlap_sets(sg, ctx):
    A = O.poisson3d(24, 20, 16)
    dA = sg.Csr(ctx, A.rp, A.ci, A.v)
    M = sg.Precond(ctx, "ras", nparts=12, overlap=2).setup(dA)
    part = M.partition()
    counts = np.bincount(part, minlength=12)
    assert counts.min() > 0 and counts.max() <= 1.03 * A.n / 12 + 8
    # deterministic
    M2 = sg.Precond(ctx, "ras", nparts=12, overlap=2).setup(dA)
    assert np.array_equal(part, M2.partition())
    # overlap sets equal the oracle's given the same partition (bit-exact)
    for s, ref in zip(M.subdomains(), O.overlap_sets(A, part, 12, 2)):
        assert np.array_equal(s, ref)
    # RAS apply equals the oracle's on that partition
    x = np.random.default_rng(6).standard_normal(A.n)
    y = dev(sg, ctx, np.zeros(A.n))
    M.apply(dev(sg, ctx, x), y)
    assert np.array_equal(y.get(), O.Precond("ras", A, 12, 2, part).apply(x))


@pytest.mark.parametrize("fine", ["twolevel_ras", "twolevel_asm"])
@pytest.mark.parametrize("coarse_solve", [0, 1])
def test_two_level_fp64(sg, ctx, fine, coarse_solve):
    A = O.poisson2d(40, 36)
    dA = sg.Csr(ctx, A.rp, A.ci, A.v)
    M = sg.Precond(ctx, fine, nparts=9, overlap=1, coarse_solve=coarse_solve).setup(dA)
    part = M.partition()
    ref = O.Precond(fine, A, 9, 1, part)
    A0, LU, piv = M.coarse()
    rA0, rLU, rpiv = ref.coarse()
    assert np.array_equal(A0, rA0)
    assert np.array_equal(piv, rpiv)
    assert np.array_equal(LU, rLU)
    x = np.random.default_rng(7).standard_normal(A.n)
    y = dev(sg, ctx, np.zeros(A.n))
    M.apply(dev(sg, ctx, x), y)
    want = ref.apply(x)
    assert np.linalg.norm(y.get() - want) <= 1e-12 * np.linalg.norm(want)


@pytest.mark.parametrize("kind", ["ilu0", "ras", "asm", "twolevel_ras"])
def test_fp32_apply_matches_oracle_mixed_mode(sg, ctx, kind):
    A = O.poisson3d(16, 16, 12)
    dA = sg.Csr(ctx, A.rp, A.ci, A.v)
    M = sg.Precond(ctx, kind, nparts=6, overlap=1, precision="fp32").setup(dA)
    part = M.partition() if kind != "ilu0" else None
    ref = O.Precond(kind, A, 6 if part is not None else 1, 1, part, precision=1)
    x = np.random.default_rng(8).standard_normal(A.n)
    y = dev(sg, ctx, np.zeros(A.n))
    M.apply(dev(sg, ctx, x), y)
    want = ref.apply(x)
    assert np.linalg.norm(y.get() - want) <= 1e-5 * np.linalg.norm(want)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_kernel_shapes_do_not_change_results(sg, precision):
    """Every SpMV block shape and sweep workgroup size (forced, and the one the
    autotuner picks) gives bit-identical SpMV and RAS/ASM/two-level applies."""
    A = O.poisson3d(40, 36, 30)
    rng = np.random.default_rng(11)
    v = A.v * rng.uniform(0.5, 1.5, A.nnz)
    x = rng.standard_normal(A.n)
    outs = []
    for mode in range(12):
        c = sg.Context(0)
        c.autotune(mode)
        dA = sg.Csr(c, A.rp, A.ci, v)
        dx = dev(sg, c, x)
        y = dev(sg, c, np.zeros(A.n))
        dA.spmv(dx, y)
        res = [y.get()]
        # coarse_solve=1: the staged shapes of at most 512 threads form the
        # coarse term inside the sweep, the others in k_coarse_inv
        for kind, kw in (("ras", {}), ("asm", {}), ("twolevel_ras", {}),
                         ("twolevel_ras", {"coarse_solve": 1})):
            M = sg.Precond(c, kind, nparts=16, overlap=1, precision=precision,
                           partition=np.arange(A.n, dtype=np.int32) * 16 // A.n, **kw).setup(dA)
            for _ in range(2):  # first call tunes in mode 1
                z = dev(sg, c, np.zeros(A.n))
                M.apply(dx, z)
                res.append(z.get())
        outs.append(res)
        del M, dA, dx, y, z
        c.close()
    assert np.array_equal(outs[0][0], O.spmv(O.Csr(A.rp, A.ci, v), x))
    for res in outs[1:]:
        for a, b in zip(res, outs[0]):
            assert np.array_equal(a, b)


def _rhs(g, A):
    from rng_ref import splitmix_uniform
    if g["rhs"] == "ones":
        return O.spmv(A, np.ones(A.n)), np.ones(A.n)
    xs = 2.0 * splitmix_uniform(int(g["rhs"].split(":")[1]), A.n) - 1.0
    return O.spmv(A, xs), xs


@pytest.mark.parametrize("idx", range(12))
def test_solve_matches_reference_golden(sg, ctx, gold, idx):
    g = gold("solve.json")[idx]
    A = O.grid(g["grid"])
    b, xs = _rhs(g, A)
    dA = sg.Csr(ctx, A.rp, A.ci, A.v)
    M = None if g["pc"] == "none" else sg.Precond(ctx, g["pc"]).setup(dA)
    db, dx = dev(sg, ctx, b), dev(sg, ctx, np.zeros(A.n))
    r = sg.solve(dA, db, dx, M, method=g["method"], restart=g["restart"],
                 max_iter=g["max_iter"], tol=g["tol"])
    assert r["converged"] == bool(g["converged"])
    assert abs(r["iterations"] - g["iterations"]) <= 1, (r["iterations"], g["iterations"])
    k = min(len(r["hist"]), len(g["hist"]))
    h, gh = r["hist"][:k], np.array(g["hist"][:k])
    # estimates agree to 1e-6 relative, or to 1e-12 of the initial residual once
    # they are far below it (tree vs sequential reductions over hundreds of steps)
    assert np.all(np.abs(h - gh) <= 1e-6 * gh + 1e-12 * gh[0])
    if "x" in g and g["converged"]:
        x = dx.get()
        assert np.linalg.norm(x - np.array(g["x"])) <= 1e-6 * np.linalg.norm(g["x"])
    if g["converged"]:
        assert r["rel_true_residual"] <= 2.0 * g["rel_true_residual"] + 1e-12


@pytest.mark.parametrize("method", ["gmres", "fgmres"])
@pytest.mark.parametrize("kind", ["ras", "twolevel_ras", "asm"])
def test_schwarz_solve_vs_oracle(sg, ctx, method, kind):
    A = O.poisson3d(24, 24, 20)
    from rng_ref import splitmix_uniform
    xs = 2.0 * splitmix_uniform(11, A.n) - 1.0
    b = O.spmv(A, xs)
    dA = sg.Csr(ctx, A.rp, A.ci, A.v)
    M = sg.Precond(ctx, kind, nparts=8, overlap=2).setup(dA)
    part = M.partition()
    ref = O.solve(method, A, b, O.Precond(kind, A, 8, 2, part), 30, 1000, 1e-9)
    db, dx = dev(sg, ctx, b), dev(sg, ctx, np.zeros(A.n))
    r = sg.solve(dA, db, dx, M, method=method, restart=30, max_iter=1000, tol=1e-9)
    assert r["converged"] and ref["converged"]
    assert abs(r["iterations"] - ref["iterations"]) <= 1
    x = dx.get()
    assert np.linalg.norm(x - ref["x"]) <= 1e-6 * np.linalg.norm(ref["x"])


@pytest.mark.parametrize("method", ["gmres", "fgmres"])
def test_rtol_mode_verifies_true_residual(sg, ctx, method):
    A = O.poisson3d(32, 32, 32)
    from rng_ref import splitmix_uniform
    xs = 2.0 * splitmix_uniform(2, A.n) - 1.0
    b = O.spmv(A, xs)
    dA = sg.Csr(ctx, A.rp, A.ci, A.v)
    M = sg.Precond(ctx, "ras", nparts=8, overlap=2).setup(dA)
    for ortho in ("mgs", "cgs2"):
        db, dx = dev(sg, ctx, b), dev(sg, ctx, np.zeros(A.n))
        r = sg.solve(dA, db, dx, M, method=method, restart=50, max_iter=2000, rtol=1e-8,
                     ortho=ortho)
        assert r["converged"], r
        x = dx.get()
        true = np.linalg.norm(b - O.spmv(A, x)) / np.linalg.norm(b)
        assert true <= 1e-8 and abs(true - r["rel_true_residual"]) <= 1e-3 * true


def test_mixed_fgmres_converges_to_fp64_accuracy(sg, ctx):
    A = O.poisson3d(32, 32, 24)
    from rng_ref import splitmix_uniform
    xs = 2.0 * splitmix_uniform(3, A.n) - 1.0
    b = O.spmv(A, xs)
    dA = sg.Csr(ctx, A.rp, A.ci, A.v)
    M = sg.Precond(ctx, "twolevel_ras", nparts=16, overlap=1, precision="fp32",
                   coarse_solve=1).setup(dA)
    db, dx = dev(sg, ctx, b), dev(sg, ctx, np.zeros(A.n))
    r = sg.solve(dA, db, dx, M, method="fgmres", restart=40, max_iter=2000, rtol=1e-8,
                 ortho="cgs2")
    assert r["converged"]
    x = dx.get()
    assert np.linalg.norm(b - O.spmv(A, x)) <= 1e-8 * np.linalg.norm(b)
    assert np.linalg.norm(x - xs) <= 1e-6 * np.linalg.norm(xs)


def test_solve_is_deterministic(sg, ctx):
    A = O.poisson2d(64, 64)
    b = O.spmv(A, np.ones(A.n))
    dA = sg.Csr(ctx, A.rp, A.ci, A.v)
    M = sg.Precond(ctx, "twolevel_ras", nparts=8, overlap=1).setup(dA)
    outs = []
    for _ in range(2):
        db, dx = dev(sg, ctx, b), dev(sg, ctx, np.zeros(A.n))
        r = sg.solve(dA, db, dx, M, method="fgmres", restart=30, tol=1e-10, ortho="cgs2")
        outs.append((r["iterations"], dx.get(), r["hist"]))
    assert outs[0][0] == outs[1][0]
    assert np.array_equal(outs[0][1], outs[1][1])
    assert np.array_equal(outs[0][2], outs[1][2])


def test_errors_are_reported(sg, ctx):
    A = O.poisson2d(8, 8)
    dA = sg.Csr(ctx, A.rp, A.ci, A.v)
    M = sg.Precond(ctx, "ras", nparts=4)
    x = dev(sg, ctx, np.zeros(A.n))
    with pytest.raises(sg.SgError) as e:
        M.apply(x, x)  # before setup
    assert e.value.status == 7
    with pytest.raises(sg.SgError) as e:
        sg.Precond(ctx, "ras", nparts=100).setup(dA)  # more parts than rows
    assert e.value.status == 4
    # singular coarse matrix: a zero block row makes A0 singular
    rp = np.arange(0, 17, dtype=np.int32)
    ci = np.arange(16, dtype=np.int32)
    v = np.ones(16)
    v[:4] = 0.0
    dZ = sg.Csr(ctx, rp, ci, v)
    part = (np.arange(16) // 4).astype(np.int32)
    with pytest.raises(sg.SgError) as e:
        sg.Precond(ctx, "twolevel_ras", nparts=4, overlap=0, partition=part).setup(dZ)
    assert e.value.status == 3


def test_tune_cache_makes_selection_deterministic(sg, tmp_path):
    """The autotuner stores its choices under a structural key; a context that
    loads the cache uses the stored shapes without timing (here a forced, edited
    entry), and results do not depend on the shape."""
    A = O.poisson3d(40, 36, 30)
    x = np.random.default_rng(21).standard_normal(A.n)
    path = tmp_path / "tune.txt"
    out = []
    for round_ in range(2):
        c = sg.Context(0)
        c.tune_cache(path)
        dA = sg.Csr(c, A.rp, A.ci, A.v)
        M = sg.Precond(c, "ras", nparts=16, overlap=1, precision="fp32",
                       partition=np.arange(A.n, dtype=np.int32) * 16 // A.n).setup(dA)
        dx, dy, dz = dev(sg, c, x), dev(sg, c, np.zeros(A.n)), dev(sg, c, np.zeros(A.n))
        dA.spmv(dx, dy)
        M.apply(dx, dz)
        out.append((M.stats()["sweep_shape"], dy.get(), dz.get()))
        if round_ == 0:
            keys = dict(l.split() for l in path.read_text().splitlines())
            assert len(keys) == 2 and any(k.startswith("spmv:") for k in keys)
            sweep_key = [k for k in keys if k.startswith("sweep:")][0]
            forced = 0 if int(keys[sweep_key]) != 0 else 2
            keys[sweep_key] = str(forced)
            path.write_text("".join(f"{k} {v}\n" for k, v in keys.items()))
        del M, dA, dx, dy, dz
        c.close()
    assert out[1][0] == forced
    assert np.array_equal(out[0][1], out[1][1]) and np.array_equal(out[0][2], out[1][2])
