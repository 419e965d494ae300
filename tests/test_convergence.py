"""Residual-history metrics of libsgamd.so (host code, no GPU needed) against
the reference's ConvergenceMetrics (src/analysis/convergence_analysis.h,
compiled under oracle/_ref): tests/golden/conv.json holds each history, the
JSONL line the reference appended and the summary it printed."""
import json

import numpy as np
import pytest

from conftest import golden


@pytest.fixture(scope="module")
def sg():
    import __graft_entry__
    return __graft_entry__.load_package()


@pytest.mark.parametrize("case", golden("conv.json"), ids=lambda c: str(len(c["hist"])))
def test_jsonl_line_matches_reference(sg, tmp_path, case):
    p = tmp_path / "c.jsonl"
    sg.convergence_write_jsonl(str(p), "m.mtx", "ras", 8, 2, case["hist"], case["converged"])
    assert p.read_text() == case["jsonl"]
    # appending keeps earlier lines
    sg.convergence_write_jsonl(str(p), "m.mtx", "ras", 8, 2, case["hist"], case["converged"])
    assert p.read_text() == case["jsonl"] * 2


@pytest.mark.parametrize("case", golden("conv.json"), ids=lambda c: str(len(c["hist"])))
def test_metrics_match_reference(sg, case):
    m = sg.convergence_metrics(case["hist"], case["converged"])
    rec = json.loads(case["jsonl"])
    for k in ("iterations", "converged", "stagnation_start"):
        assert m[k] == rec[k]
    for k in ("initial_residual", "final_residual", "asymptotic_rate"):
        assert float("%.6g" % m[k]) == rec[k]
    # the summary lines the reference printed carry the same numbers
    assert ("Stagnation detected at iter %d" % m["stagnation_start"] in case["summary"]) == \
        (m["stagnation_start"] >= 0)


def test_metrics_closed_forms(sg):
    # geometric decay r_k = q^(k-1): rate_k = q^((k-1)/k), mean of the last 20
    q, n = 0.5, 60
    h = q ** np.arange(n)
    m = sg.convergence_metrics(h, True)
    k = np.arange(1, n + 1)
    rates = q ** ((k - 1) / k)
    assert np.allclose(m["rates"], rates, rtol=1e-15)
    assert np.isclose(m["asymptotic_rate"], rates[-20:].mean(), rtol=1e-15)
    assert m["stagnation_start"] == -1
    # flat tail: the first window of 20 steps reducing by < 0.1% starts at 11
    h = np.concatenate([np.geomspace(1, 1e-3, 30), np.full(30, 1e-3)])
    m = sg.convergence_metrics(h, False)
    w = next(i for i in range(20, 60) if (h[i - 20] - h[i]) / h[i - 20] < 1e-3)
    assert m["stagnation_start"] == w - 19
    # empty history: everything zero, no stagnation
    m = sg.convergence_metrics([], True)
    assert (m["iterations"], m["initial_residual"], m["stagnation_start"]) == (0, 0.0, -1)
