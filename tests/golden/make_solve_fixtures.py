"""Write the oracle's solves of the three dense workloads to ``oracle_solves.json``.

The oracle (oracle/sdpx_oracle.py, the Float64 restatement of the reference's
dense interior-point path) solves each seeded problem from the strictly
feasible start X0 = I, y0, Z0 = I, to the gap and residual tolerance named:

* ``dense_n50_m200``   -- one PSD block n = 50, m = 200 (the reference's CPU case), seed 0, 1e-8
* ``dense_n400_m4000`` -- n = 400, m = 4000 (the FP64 GPU workload), seed 0, 1e-8
* ``dense_n500_m6000`` -- n = 500, m = 6000 (the mixed-precision workload), seed 0, 1e-8

and records status, iterations, objectives, gap and the per-iteration history,
so the GPU tests and the benchmark compare against the oracle without
re-solving on the GPU host.  The n = 500 case needs about 40 GB of host memory
and a few minutes per iteration on 8 cores.

Run: ``python tests/golden/make_solve_fixtures.py [name ...]`` (all by default;
existing entries for other names are kept).
"""

import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path[:0] = [str(ROOT / "sdpx.jl_amd"), str(ROOT / "oracle")]

import sdpx_oracle as O  # noqa: E402
from sdpx_mi355x import generators as G  # noqa: E402

OUT = Path(__file__).resolve().parent / "oracle_solves.json"
CASES = {"dense_n50_m200": (50, 200), "dense_n400_m4000": (400, 4000), "dense_n500_m6000": (500, 6000)}


def solve_case(name: str) -> dict:
    k, m = CASES[name]
    tol = 1e-8
    g = G.dense_feasible(k, m, seed=0)
    prob = O.ingest(g.c, [g.A_blocks()], [g.C])
    x0 = g.x0
    del g  # the oracle holds its own copy of the panel
    o = O.Options(eps_gap=tol, eps_primal=tol, eps_dual=tol)
    t0 = time.perf_counter()
    r = O.solve(prob, o, x0=x0, X0=[np.eye(k)], Y0=[np.eye(k)])
    dt = time.perf_counter() - t0
    print(f"{name}: {r.status} iterations={r.iterations} pobj={r.pobj:.15e} gap={r.gap_rel:.2e} {dt:.1f}s",
          flush=True)
    return {"n": k, "m": m, "seed": 0, "tol": tol, "status": r.status, "iterations": r.iterations,
            "pobj": r.pobj, "dobj": r.dobj, "gap_rel": r.gap_rel, "p_res": r.p_res, "d_res": r.d_res,
            "seconds": round(dt, 1), "blas_threads": os.environ.get("OMP_NUM_THREADS", str(os.cpu_count())),
            "history": [{key: (float(v) if isinstance(v, (float, np.floating)) else v) for key, v in row.items()}
                        for row in r.history]}


def main(names):
    data = json.loads(OUT.read_text()) if OUT.exists() else {}
    for name in names:
        data[name] = solve_case(name)
        OUT.write_text(json.dumps(data, indent=1) + "\n")


if __name__ == "__main__":
    main(sys.argv[1:] or list(CASES))
