"""Write the golden fixtures under tests/golden/ from the reference's own code.

Runs oracle/_ref/ref_harness (built by ``make -C oracle ref`` from the
reference's METIS-free sources: core/gmres.cpp, core/fgmres.cpp,
precond/ilu0.h, precond/jacobi.h, precond/coarse_solve.h, mesh/poisson.h,
io/metrics_logger.h, core/sparse_matrix.h) and stores inputs and outputs as
small JSON fixtures.  Subdomain index sets for the RAS/ASM fixtures are built
here by a brute-force breadth-first search over the dense adjacency of the
grid, independent of the oracle.

Usage:  python tests/golden/make_golden.py [bicgstab analysis conv mm ilut]
        (needs /root/reference; with names, only those fixtures are rewritten)
"""
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")


def run(*args, stdin=None):
    out = subprocess.run([HARNESS, *map(str, args)], input=stdin, capture_output=True,
                         text=True, check=True)
    return json.loads(out.stdout)


def dump(name, obj):
    with open(os.path.join(HERE, name), "w") as f:
        json.dump(obj, f, separators=(",", ":"))


def grid_adjacency(spec):
    """Dense boolean adjacency (pattern incl. diagonal) of the Poisson grid."""
    dims = [int(t) for t in spec.split(":")[1:]]
    if len(dims) == 2:
        nx, ny = dims
        n = nx * ny
        A = np.zeros((n, n), dtype=bool)
        for iy in range(ny):
            for ix in range(nx):
                i = iy * nx + ix
                A[i, i] = True
                for dx, dy in ((1, 0), (-1, 0), (0, 1), (0, -1)):
                    jx, jy = ix + dx, iy + dy
                    if 0 <= jx < nx and 0 <= jy < ny:
                        A[i, jy * nx + jx] = True
        return A
    nx, ny, nz = dims
    n = nx * ny * nz
    A = np.zeros((n, n), dtype=bool)
    for iz in range(nz):
        for iy in range(ny):
            for ix in range(nx):
                i = (iz * ny + iy) * nx + ix
                A[i, i] = True
                for d in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
                    jx, jy, jz = ix + d[0], iy + d[1], iz + d[2]
                    if 0 <= jx < nx and 0 <= jy < ny and 0 <= jz < nz:
                        A[i, (jz * ny + jy) * nx + jx] = True
    return A


def bfs_sets(adj, part, nparts, overlap):
    sets = []
    for s in range(nparts):
        cur = part == s
        for _ in range(overlap):
            cur = cur | adj[cur].any(axis=0)
        sets.append(np.flatnonzero(cur))
    return sets


def blocks_case(spec, part, nparts, overlap, seed):
    adj = grid_adjacency(spec)
    n = adj.shape[0]
    sets = bfs_sets(adj, part, nparts, overlap)
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, n)
    lines = [f"{spec} {nparts}"]
    for s, d in enumerate(sets):
        lines.append(str(len(d)))
        lines.append(" ".join(map(str, d)))
        lines.append(" ".join("1" if part[i] == s else "0" for i in d))
    lines.append(" ".join(repr(float(t)) for t in x))
    out = run("blocks", stdin="\n".join(lines) + "\n")
    return {"grid": spec, "nparts": nparts, "overlap": overlap, "part": part.tolist(),
            "sets": [d.tolist() for d in sets], "x": x.tolist(), "ras": out["ras"],
            "asm": out["asm"]}


# Matrix Market inputs for the reader fixtures: (name, text).  They cover the
# reference's header handling (flags from any '%' line), symmetric mirroring,
# pattern files, unsorted and duplicate entries, CRLF line ends, missing and
# unreadable values, number formats, rectangular shapes and its two error
# messages.
MM_CASES = [
    ("general", "%%MatrixMarket matrix coordinate real general\n% a comment\n4 4 7\n"
                "4 4 2.5\n1 1 4\n2 1 -1\n1 2 -1e-3\n3 3 1.25E+2\n2 2 4\n3 4 .5\n"),
    ("symmetric", "%%MatrixMarket matrix coordinate real symmetric\n5 5 8\n1 1 2\n2 1 -1\n"
                  "2 2 2\n3 2 -1\n3 3 2\n5 3 -0.25\n4 4 2\n5 5 2\n"),
    ("pattern_sym", "%%MatrixMarket matrix coordinate pattern symmetric\n4 4 5\n1 1\n2 1\n"
                    "3 3\n4 2\n4 4\n"),
    ("duplicates", "%%MatrixMarket matrix coordinate real general\n3 3 6\n1 1 1\n2 2 2\n"
                   "1 1 3\n3 3 5\n2 2 7\n1 3 1\n"),
    ("comment_flag", "%%MatrixMarket matrix coordinate real general\n"
                     "% this matrix is not symmetric\n3 3 3\n1 1 1\n2 1 5\n3 3 1\n"),
    ("crlf", "%%MatrixMarket matrix coordinate real general\r\n3 3 3\r\n1 1 1.5\r\n"
             "2 3 -2\r\n3 3 4\r\n"),
    ("integer", "%%MatrixMarket matrix coordinate integer general\n2 2 3\n1 1 3\n1 2 -7\n"
                "2 2 +9\n"),
    ("missing_value", "%%MatrixMarket matrix coordinate real general\n2 2 3\n1 1\n1 2 abc\n"
                      "2 2 1e\n"),
    ("numbers", "%%MatrixMarket matrix coordinate real general\n3 3 6\n1 1 1e400\n"
                "1 2 -1e400\n2 2 1e-310\n2 3 -0.0\n3 3 123456789012345678901234567890\n"
                "3 1 0.1234567890123456789 trailing tokens\n"),
    ("skew", "%%MatrixMarket matrix coordinate real skew-symmetric\n3 3 2\n2 1 3\n3 2 -4\n"),
    ("rect", "%%MatrixMarket matrix coordinate real general\n   2   5  4  \n2 5 1\n1 4 2\n"
             "1 1 3\n2 2 4\n"),
    ("premature", "%%MatrixMarket matrix coordinate real general\n3 3 4\n1 1 1\n2 2 1\n"),
    ("empty_rows", "%%MatrixMarket matrix coordinate real general\n5 5 2\n4 2 1\n4 1 2\n"),
]


def gen_bicgstab():
    """BiCGSTAB (src/core/bicgstab.h): relative tol, right preconditioning."""
    cases = [
        ("bicgstab", "ilu0", "2d:30:30", "ones", 0, 1000, 1e-10),
        ("bicgstab", "none", "2d:30:30", "ones", 0, 1000, 1e-10),
        ("bicgstab", "jacobi", "2d:40:40", "uniform:2", 0, 1000, 1e-10),
        ("bicgstab", "ilu0", "3d:12:12:12", "uniform:3", 0, 1000, 1e-10),
        ("bicgstab", "ilu0", "2d:100:100", "ones", 0, 1000, 1e-10),
        ("bicgstab", "none", "2d:30:30", "uniform:4", 0, 9, 1e-10),    # iteration bound
        ("bicgstab", "jacobi", "2d:20:20", "ones", 0, 1000, 1e-3),     # loose: half-step exit likely
        ("bicgstab", "ilu0", "rnd:300:5", "uniform:6", 0, 1000, 1e-10),  # nonsymmetric
    ]
    out = []
    for c in cases:
        r = run("solve", *c)
        r.update(dict(method=c[0], pc=c[1], grid=c[2], rhs=c[3], max_iter=c[5], tol_arg=c[6]))
        out.append(r)
    dump("bicgstab.json", out)


def gen_analysis():
    """estimate_condition / estimate_spectral_radius (src/analysis/*.h)."""
    cases = [("none", "2d:20:20", 80, 300), ("jacobi", "2d:20:20", 80, 300),
             ("ilu0", "2d:30:30", 80, 300), ("ilu0", "3d:8:8:8", 80, 300),
             ("none", "2d:6:6", 80, 300), ("ilu0", "rnd:200:7", 40, 200),
             ("jacobi", "2d:3:1", 80, 50)]
    out = []
    for pc, spec, k, pit in cases:
        r = run("analyze", pc, spec, k, pit)
        r.update(dict(pc=pc, grid=spec, k=k, power_iters_max=pit))
        out.append(r)
    dump("analysis.json", out)


def gen_conv():
    """ConvergenceMetrics (src/analysis/convergence_analysis.h): JSONL line and summary."""
    rng = np.random.default_rng(21)
    hists = [
        (np.geomspace(1.0, 1e-9, 40) * (1 + 0.1 * rng.standard_normal(40))).tolist(),
        ([1.0] + [0.5] * 30 + [0.4]),  # stagnation
        [3.5],
        [],
        (np.geomspace(2e3, 1e-3, 25)).tolist(),
    ]
    out = []
    import tempfile
    for i, h in enumerate(hists):
        conv = i % 2
        with tempfile.TemporaryDirectory() as td:
            path = os.path.join(td, "c.jsonl")
            txt = f"{len(h)}\n" + " ".join(repr(float(t)) for t in h) + "\n"
            so = subprocess.run([HARNESS, "conv", path, str(conv)], input=txt,
                                capture_output=True, text=True, check=True).stdout
            line = open(path).read()
        out.append({"hist": [float(t) for t in h], "converged": conv, "jsonl": line,
                    "summary": so})
    dump("conv.json", out)


def gen_mm():
    """read_matrix_market / write_matrix_market (src/io/matrix_market.h)."""
    import tempfile
    reads = []
    with tempfile.TemporaryDirectory() as td:
        for name, text in MM_CASES:
            path = os.path.join(td, name + ".mtx")
            with open(path, "w", newline="") as f:
                f.write(text)
            r = run("mm", path)
            if "error" in r:
                r["error"] = r["error"].replace(td + os.sep, "")
            reads.append({"name": name, "text": text, "out": r})
        writes = []
        for spec in ("2d:3:3", "rnd:12:3"):
            path = os.path.join(td, "w.mtx")
            m = run("mmw", spec, path)
            writes.append({"grid": spec, "matrix": m, "text": open(path).read()})
    dump("mm.json", {"reads": reads, "writes": writes})


def gen_ilut():
    """ILUTPrecond (src/precond/ilut.h): factor sizes, one apply, GMRES / BiCGSTAB counts."""
    applies = []
    # every case keeps each row's index list within the 2 * fill * (nnz / n)
    # entries the reference reserves: beyond that its elimination loop reads a
    # reallocated vector (undefined; rnd:150:8 with drop 5e-2, fill 1 crashes)
    for spec, drop, fill in (("2d:12:12", 1e-4, 10), ("2d:20:20", 1e-2, 2), ("3d:6:6:6", 1e-3, 5),
                             ("rnd:150:8", 1e-4, 10), ("rnd:150:8", 5e-2, 3), ("2d:30:30", 0.0, 1),
                             ("3d:8:8:8", 1e-4, 10), ("2d:30:30", 0.0, 3)):
        r = run("ilut", spec, drop, fill)
        r.update(dict(grid=spec, drop_tol=drop, fill_factor=fill))
        applies.append(r)
    solves = []
    for c in (("gmres", "ilut", "2d:60:60", "ones", 30, 1000, 1e-10),
              ("fgmres", "ilut", "3d:12:12:12", "uniform:2", 20, 1000, 1e-10),
              ("bicgstab", "ilut", "2d:60:60", "uniform:3", 0, 1000, 1e-10)):
        r = run("solve", *c)
        r.update(dict(method=c[0], pc=c[1], grid=c[2], rhs=c[3], restart=c[4], max_iter=c[5],
                      tol_arg=c[6]))
        solves.append(r)
    dump("ilut.json", {"applies": applies, "solves": solves})


def main():
    if not os.path.exists(HARNESS):
        sys.exit("build the harness first: make -C oracle ref")
    only = set(sys.argv[1:])
    extra = {"bicgstab": gen_bicgstab, "analysis": gen_analysis, "conv": gen_conv, "mm": gen_mm,
             "ilut": gen_ilut}
    if only:
        for k in sorted(only):
            extra[k]()
        print("fixtures written to", HERE)
        return
    for f in extra.values():
        f()

    # ILU(0) factors and one apply (src/precond/ilu0.h).
    ilu = {spec: run("ilu0", spec) for spec in ("2d:8:8", "2d:16:16", "3d:6:6:6", "3d:5:4:3")}
    dump("ilu0.json", ilu)

    # GMRES / FGMRES iteration counts and residual histories.
    cases = [
        ("gmres", "ilu0", "2d:100:100", "ones", 30, 1000, 1e-10),
        ("gmres", "ilu0", "2d:150:150", "ones", 30, 1000, 1e-10),
        ("gmres", "ilu0", "2d:256:256", "uniform:1", 30, 2000, -1e-8),
        ("fgmres", "ilu0", "2d:256:256", "uniform:1", 30, 2000, -1e-8),
        ("gmres", "none", "2d:20:20", "ones", 30, 1000, 1e-10),
        ("gmres", "jacobi", "2d:20:20", "uniform:3", 25, 1000, 1e-10),
        ("fgmres", "none", "2d:20:20", "uniform:4", 30, 1000, 1e-10),
        ("fgmres", "jacobi", "3d:8:8:8", "uniform:5", 20, 1000, 1e-10),
        ("gmres", "ilu0", "3d:16:16:16", "uniform:6", 50, 2000, -1e-8),
        ("fgmres", "ilu0", "3d:16:16:16", "uniform:6", 50, 2000, -1e-8),
        ("gmres", "ilu0", "2d:5:5", "ones", 25, 1000, 1e-10),
        ("gmres", "none", "2d:30:30", "ones", 10, 7, 1e-14),  # max_iter overshoot
    ]
    solves = []
    for c in cases:
        r = run("solve", *c)
        r.update(dict(method=c[0], pc=c[1], grid=c[2], rhs=c[3], restart=c[4],
                      max_iter=c[5], tol_arg=c[6]))
        solves.append(r)
    dump("solve.json", solves)

    # RAS / ASM applies on reference extract_submatrix + ILU0Precond blocks.
    part20 = np.array([(0 if (i % 20) < 10 else 1) + (0 if (i // 20) < 10 else 2)
                       for i in range(400)])
    part3d = np.array([((i % 6) // 3) + 2 * (((i // 36) % 6) // 3) for i in range(216)])
    rng = np.random.default_rng(11)
    part_rand = rng.integers(0, 5, 144)
    blocks = [
        blocks_case("2d:20:20", part20, 4, 1, 1),
        blocks_case("2d:20:20", part20, 4, 2, 2),
        blocks_case("2d:20:20", part20, 4, 0, 3),
        blocks_case("3d:6:6:6", part3d, 4, 1, 4),
        blocks_case("2d:12:12", part_rand, 5, 2, 5),
    ]
    dump("blocks.json", blocks)

    # Dense coarse LU with partial pivoting (src/precond/coarse_solve.h).
    dump("lu.json", {"n16": run("lu", 16, 5, 0), "n40": run("lu", 40, 6, 0),
                     "singular": run("lu", 8, 9, 1)})

    # JSONL records (src/io/metrics_logger.h).
    lines = subprocess.run([HARNESS, "metrics"], capture_output=True, text=True,
                           check=True).stdout.splitlines()
    dump("metrics.json", lines)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
