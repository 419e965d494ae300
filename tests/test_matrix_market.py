"""Matrix Market reader / writer of libsgamd.so (host code, no GPU needed).

Default mode is pinned against the reference's own read_matrix_market /
write_matrix_market (src/io/matrix_market.h:14-103), compiled under
oracle/_ref: tests/golden/mm.json holds each input file's text and the CSR
(or error message) the reference produced for it, and the exact bytes it wrote
for two matrices.  The fixes behind flags (summing duplicates, strict header
and line handling) have no reference counterpart; they are checked against
hand-derived expectations and dense sums.
"""
import os

import numpy as np
import pytest

from conftest import golden


@pytest.fixture(scope="module")
def sg():
    import __graft_entry__
    return __graft_entry__.load_package()


def _write(tmp_path, name, text):
    p = tmp_path / (name + ".mtx")
    with open(p, "w", newline="") as f:
        f.write(text)
    return str(p)


def _rows(rp, ci, v):
    return [sorted(zip(ci[rp[i]:rp[i + 1]].tolist(), v[rp[i]:rp[i + 1]].tolist()))
            for i in range(len(rp) - 1)]


@pytest.mark.parametrize("case", golden("mm.json")["reads"], ids=lambda c: c["name"])
def test_reader_matches_reference(sg, tmp_path, case):
    path = _write(tmp_path, case["name"], case["text"])
    ref = case["out"]
    if "error" in ref:
        with pytest.raises(sg.SgError) as e:
            sg.read_matrix_market(path)
        assert str(ref["error"]) in str(e.value)
        return
    n, nc, rp, ci, v = sg.read_matrix_market(path)
    assert (n, nc) == (ref["nrows"], ref["ncols"])
    assert rp.tolist() == ref["row_ptr"]
    assert ci.tolist() == ref["col_idx"]
    # duplicates keep file order here; the reference's unstable sort leaves
    # their order unspecified, so compare each row as a multiset as well
    assert _rows(rp, ci, v) == _rows(np.array(ref["row_ptr"]), np.array(ref["col_idx"]),
                                     np.array(ref["vals"], dtype=float))
    if case["name"] != "duplicates":
        assert v.tolist() == [float(t) for t in ref["vals"]]


def test_reader_missing_file_message(sg, tmp_path):
    path = str(tmp_path / "nope.mtx")
    with pytest.raises(sg.SgError) as e:
        sg.read_matrix_market(path)
    assert "Cannot open file: " + path in str(e.value)


@pytest.mark.parametrize("w", golden("mm.json")["writes"], ids=lambda w: w["grid"])
def test_writer_matches_reference_bytes(sg, tmp_path, w):
    m = w["matrix"]
    p = str(tmp_path / "out.mtx")
    sg.write_matrix_market(p, m["row_ptr"], m["col_idx"], np.array(m["vals"], dtype=float),
                           ncols=m["ncols"])
    assert open(p).read() == w["text"]


def test_round_trip_17_digits(sg, tmp_path):
    rng = np.random.default_rng(3)
    n = 300
    dense = (rng.random((n, n)) < 0.02) * rng.standard_normal((n, n)) * 10.0 ** rng.integers(
        -20, 20, (n, n))
    np.fill_diagonal(dense, 1.0 + rng.random(n))
    rows, cols = np.nonzero(dense)
    rp = np.searchsorted(rows, np.arange(n + 1)).astype(np.int32)
    p = str(tmp_path / "rt.mtx")
    sg.write_matrix_market(p, rp, cols, dense[rows, cols], digits=17)
    n2, nc2, rp2, ci2, v2 = sg.read_matrix_market(p)
    assert n2 == nc2 == n
    assert np.array_equal(rp2, rp) and np.array_equal(ci2, cols)
    assert np.array_equal(v2, dense[rows, cols])  # bit for bit


def test_sum_duplicates_flag(sg, tmp_path):
    text = next(c["text"] for c in golden("mm.json")["reads"] if c["name"] == "duplicates")
    path = _write(tmp_path, "dup", text)
    n, nc, rp, ci, v = sg.read_matrix_market(path, sum_duplicates=True)
    # (1,1): 1 + 3, (2,2): 2 + 7, (3,3): 5, (1,3): 1
    assert rp.tolist() == [0, 2, 3, 4]
    assert ci.tolist() == [0, 2, 1, 2]
    assert v.tolist() == [4.0, 1.0, 9.0, 5.0]


def test_sum_duplicates_equals_dense_sum(sg, tmp_path):
    rng = np.random.default_rng(5)
    n, k = 40, 600
    r = rng.integers(0, n, k)
    c = rng.integers(0, n, k)
    val = rng.standard_normal(k)
    lines = [f"{n} {n} {k}"] + [f"{a + 1} {b + 1} {float(x)!r}" for a, b, x in zip(r, c, val)]
    path = _write(tmp_path, "rand", "%%MatrixMarket matrix coordinate real general\n" +
                  "\n".join(lines) + "\n")
    _, _, rp, ci, v = sg.read_matrix_market(path, sum_duplicates=True)
    D = np.zeros((n, n))
    np.add.at(D, (r, c), val)
    got = np.zeros((n, n))
    for i in range(n):
        cols = ci[rp[i]:rp[i + 1]]
        assert np.all(np.diff(cols) > 0)  # merged and sorted
        got[i, cols] = v[rp[i]:rp[i + 1]]
    assert np.allclose(got, D, rtol=0, atol=1e-14)
    # without the flag every entry is kept
    _, _, rp0, _, _ = sg.read_matrix_market(path)
    assert rp0[-1] == k


def test_strict_mode(sg, tmp_path):
    cases = {c["name"]: c["text"] for c in golden("mm.json")["reads"]}
    # the word "symmetric" in a comment does not mirror
    _, _, rp, ci, v = sg.read_matrix_market(_write(tmp_path, "cf", cases["comment_flag"]),
                                            strict=True)
    assert rp.tolist() == [0, 1, 2, 3] and ci.tolist() == [0, 0, 2]
    # skew-symmetric mirrors with the opposite sign
    _, _, rp, ci, v = sg.read_matrix_market(_write(tmp_path, "sk", cases["skew"]), strict=True)
    assert ci.tolist() == [1, 0, 2, 1] and v.tolist() == [-3.0, 3.0, 4.0, -4.0]
    # blank and comment lines among the entries are skipped
    t = ("%%MatrixMarket matrix coordinate real general\n\n% c\n2 2 2\n\n1 1 5\n% mid\n"
         "2 2 6\n\n")
    _, _, rp, ci, v = sg.read_matrix_market(_write(tmp_path, "bl", t), strict=True)
    assert v.tolist() == [5.0, 6.0]
    with pytest.raises(sg.SgError):  # the reference's parse of this is undefined
        sg.read_matrix_market(_write(tmp_path, "bl2", t))
    bad = {
        "nobanner": "2 2 1\n1 1 1\n",
        "array": "%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n",
        "complex": "%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1 0\n",
        "badval": "%%MatrixMarket matrix coordinate real general\n1 1 1\n1 1 x\n",
        "missing": "%%MatrixMarket matrix coordinate real general\n1 1 1\n1 1\n",
        "extra": "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1 1\n2 2 1\n",
    }
    for name, text in bad.items():
        with pytest.raises(sg.SgError):
            sg.read_matrix_market(_write(tmp_path, name, text), strict=True)
    # the default mode reads the last three as the reference does
    assert sg.read_matrix_market(_write(tmp_path, "bv", bad["badval"]))[4].tolist() == [0.0]
    assert sg.read_matrix_market(_write(tmp_path, "mv", bad["missing"]))[4].tolist() == [1.0]
    assert sg.read_matrix_market(_write(tmp_path, "ex", bad["extra"]))[4].tolist() == [1.0]


@pytest.mark.parametrize("text", [
    "%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1\n",   # row out of range
    "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 0 1\n",   # column 0
    "%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1\n\n",  # blank entry line
    "%%MatrixMarket matrix coordinate real general\n\n2 2 1\n1 1 1\n",  # blank size line
    "%%MatrixMarket matrix coordinate real general\nx y z\n",
    "%%MatrixMarket matrix coordinate real symmetric\n2 3 1\n3 1 1\n",  # mirror outside
])
def test_undefined_reference_inputs_fail_cleanly(sg, tmp_path, text):
    with pytest.raises(sg.SgError):
        sg.read_matrix_market(_write(tmp_path, "u", text))
