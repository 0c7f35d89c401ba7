"""The mixed-precision configuration at its own size (one PSD block n = 500, m = 6000, rel gap 1e-8).

Each precision route -- FP64, FP32-assembled Schur Gram and BF16x3-assembled Schur Gram, both refined
against the matrix-free FP64 operator, with the static condition cutoff ("auto") and as the measured
trial ("on") -- is compared with the oracle's own solve of the same problem
(tests/golden/oracle_solves.json, written by make_solve_fixtures.py): same status, iterations within
2, objective within 1e-8 relative; the reduced-precision routes must have run some iterations in
reduced precision, and the three GPU objectives agree with each other to 1e-8.
"""

import json
from pathlib import Path

import pytest

from sdpx_mi355x import MI355XDenseBackend, SolverOptions, generators, ingest, solve

pytestmark = pytest.mark.gpu

K, M = 500, 6000
FIXTURES = Path(__file__).resolve().parent / "golden" / "oracle_solves.json"


def test_mixed_fullsize_matches_oracle(gpu_lib):
    g = generators.dense_feasible(K, M, seed=0)
    prob = ingest(g.c, [g.A_blocks()], [g.C])
    ref = json.loads(FIXTURES.read_text())["dense_n500_m6000"]
    tol = 1e-8
    b = MI355XDenseBackend(prob)
    try:
        b.autotune()
        results = {}
        for name, prec, mode in (("fp64", "fp64", "auto"), ("mixed", "mixed", "auto"),
                                 ("mixed_bf16x3", "mixed_bf16x3", "auto"), ("mixed_on", "mixed", "on"),
                                 ("mixed_bf16x3_on", "mixed_bf16x3", "on")):
            o = SolverOptions(eps_gap=tol, eps_primal=tol, eps_dual=tol, precision=prec, mixed_mode=mode)
            r = solve(prob, o, x0=g.x0, X0=[g.X0], Y0=[g.Y0], backend=b)
            results[name] = r
            prec = name
            print(f"{prec}: {r.status} iters={r.iterations} reduced={r.fp32_iterations} pobj={r.pobj:.12e} "
                  f"| oracle {ref['status']} iters={ref['iterations']} pobj={ref['pobj']:.12e}", flush=True)
    finally:
        b.close()
    for prec, r in results.items():
        assert r.status == ref["status"] == "Optimal" and r.gap_rel <= tol, prec
        assert abs(r.iterations - ref["iterations"]) <= 2, prec
        assert abs(r.pobj - ref["pobj"]) <= 1e-8 * max(1.0, abs(ref["pobj"])), prec
    assert all(r.fp32_iterations > 0 for name, r in results.items() if name != "fp64")
    p64 = results["fp64"].pobj
    assert all(abs(r.pobj - p64) <= 1e-8 * max(1.0, abs(p64)) for r in results.values())
