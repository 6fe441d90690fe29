"""Writes workloads.npz: the FP64 oracle's outputs on the benchmark workloads too large to solve with the
oracle inside a test run (minutes of CPU time each).

  pairwise2048   2048 x 2048, r = exp(2 z) float32, 4 neighbours, 16 focal cells: all 120 pairs'
                 resistances and the cumulative current map (summary)
  adjoint1024    1024 x 1024 lognormal float32, 8 neighbours, 32 pairs: J, the resistances and dJ/dr
                 (summary)
  alltoone1024   a 1024 x 1024 crop of the 8192 x 8192 large-graph map (10 % NoData), 4 neighbours,
                 64 focal cells, all-to-one: the cumulative current map (summary)

Each entry records the SHA-256 of the inputs it was made from, so a test can tell a changed input
generator from a wrong result. Run from the repository root (about 10 minutes on 8 cores):
``python tests/golden/make_workload_vectors.py [pairwise2048 adjoint1024 alltoone1024]``.
"""

from __future__ import annotations

import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parent.parent))

import workloads as W  # noqa: E402
from oracle import oracle as O  # noqa: E402
from oracle_core import OracleCore  # noqa: E402


def pairwise2048(out: dict) -> None:
    g, cells = W.pairwise_2048()
    n = cells.size
    r = O.cs_pairwise(g, None, cells, np.zeros((n, n), np.uint8), False, True, {"cum": True})
    out["pairwise2048/input_sha"] = np.array(W.digest(g, cells))
    out["pairwise2048/R"] = r["resistances"]
    for k, v in W.summary(r["cum"], 2048).items():
        out[f"pairwise2048/cum/{k}"] = v
    out["pairwise2048/cum_focal"] = r["cum"].ravel()[cells]


def adjoint1024(out: dict) -> None:
    r, src, dst = W.adjoint_1024()
    J, R, grad = O.adjoint(r, src, dst, False)
    out["adjoint1024/input_sha"] = np.array(W.digest(r, src, dst))
    out["adjoint1024/J"] = np.array(J)
    out["adjoint1024/R"] = R
    for k, v in W.summary(grad, 1024).items():
        out[f"adjoint1024/grad/{k}"] = v


def alltoone1024(out: dict) -> None:
    r, pm = W.large_graph_crop()
    res = W.all_to_one(r, pm, OracleCore(O))
    out["alltoone1024/input_sha"] = np.array(W.digest(r, pm))
    for k, v in W.summary(res.cum_current, 1105).items():
        out[f"alltoone1024/cum/{k}"] = v


def main(names) -> None:
    out = dict(np.load(W.GOLDEN)) if W.GOLDEN.exists() else {}
    for name in names:
        t0 = time.perf_counter()
        for k in [k for k in out if k.startswith(name + "/")]:
            del out[k]
        globals()[name](out)
        print(f"{name}: {time.perf_counter() - t0:.1f} s", flush=True)
        np.savez_compressed(W.GOLDEN, **out)


if __name__ == "__main__":
    main(sys.argv[1:] or ["pairwise2048", "adjoint1024", "alltoone1024"])
