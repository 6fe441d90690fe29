"""``core()`` answered by the FP64 oracle: the binding's functions over the oracle's modes.

Tests swap it in for the HIP library to check the Python layer on the CPU.
"""

from __future__ import annotations

import numpy as np

from landschaft_amd.circuit import _backend


class OracleCore:
    """``core()`` answered by the oracle, outputs converted to the binding's per-node form."""

    def __init__(self, oracle):
        self.o = oracle

    def circuit_advanced(self, g, s, gr, four, n_threads=0):
        return self.o.advanced(g, s, gr, four)

    @staticmethod
    def _per_node(nm, cells):
        on = nm >= 0
        out = np.zeros(int(nm.max()) + 1 if on.any() else 0)
        out[nm[on]] = cells[on]
        return out

    def circuitscape_advanced(self, g, poly, s, gr, avg, four, policy, n_threads=0):
        r = self.o.cs_advanced(g, poly, s, gr, avg, four, policy)
        nm = r["nodemap"]
        return {"nodemap": nm, "voltage": self._per_node(nm, r["voltage"]),
                "current": self._per_node(nm, r["current"]), "solved": r["solved"],
                "n_components": r["n_components"]}

    def circuitscape_pairwise(self, g, poly, cells, ex, avg, four, opts, emit=None, n_threads=0):
        if opts.get("region_ids") is not None:
            # focal regions: the per-pair graphs of the binding, each solved by the oracle
            return _backend.pairwise_regions(self.o.cs_pairwise, g, poly, cells, opts["region_ids"], ex, avg, four,
                                             opts, emit)
        return self.o.cs_pairwise(g, poly, cells, ex, avg, four, opts, emit)

    def circuitscape_multi_advanced_sum(self, g, poly, avg, four, problems, policy, cum, max_map, focal_zero=False,
                                        n_threads=0):
        out = self.circuitscape_multi_advanced(g, poly, avg, four, problems, policy)
        nm = out["nodemap"]
        on = nm >= 0
        vcheck = np.zeros(len(problems))
        for p, (q, r) in enumerate(zip(problems, out["results"])):
            cur = np.zeros(nm.shape)
            cur[on] = r["current"][nm[on]]
            if focal_zero:
                cur[np.isin(nm, r["active"])] = 0.0
            cum += cur
            if max_map is not None:
                np.maximum(max_map, cur, out=max_map)
            node = nm.flat[q["check_cell"]]
            vcheck[p] = r["voltage"][node] if node >= 0 else 0.0
        return {"nodemap": nm, "check_voltage": vcheck, "solved": np.array([r["solved"] for r in out["results"]])}

    def circuitscape_multi_advanced(self, g, poly, avg, four, problems, policy, n_threads=0):
        r = self.o.cs_multi_advanced(g, poly, avg, four, problems, policy)
        nm = r["nodemap"]
        on = nm >= 0
        results = []
        for p in range(len(problems)):
            results.append({"voltage": self._per_node(nm, r["voltage"][p]),
                            "current": self._per_node(nm, r["current"][p]),
                            "active": np.unique(nm[r["active"][p] & on]).astype(np.int32),
                            "solved": bool(r["solved"][p])})
        return {"nodemap": nm, "results": results}

    @staticmethod
    def ascii_grid_rows(values):
        # grid text is host code of the library (no device work): the writer is not under test here
        return _backend.ascii_grid_rows(values)

    def network_advanced(self, a, b, w, m, src, gnd, policy, n_threads=0):
        from oracle import network as ON

        return ON.network_advanced(a, b, w, m, src, gnd, policy)

    def network_pairwise(self, a, b, w, m, points, ex, opts, emit=None, n_threads=0):
        from oracle import network as ON

        return ON.network_pairwise(a, b, w, m, points, ex, opts, emit)
