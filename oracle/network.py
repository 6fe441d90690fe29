"""Circuitscape's network mode restated in numpy — TEST INFRASTRUCTURE ONLY.

Only tests/ use this module, as the checker; the product path
(landschaft_amd) never imports it. Dense solves, so graphs of a few thousand
nodes at most. Each function restates the reference entry point it names
(paths relative to the reference repository):

  network_graph ....... src/cpp/circuit/circuitscape.cpp:265-338
  node_currents ....... src/cpp/circuit/circuitscape.cpp:36-73 (graph_node_currents)
  branch_currents ..... src/cpp/circuit/circuitscape.cpp:340-362 (network_currents)
  network_advanced .... src/cpp/circuit/python.cpp:310-346 (raster_advanced, circuitscape.cpp:393-468)
  network_pairwise .... src/cpp/circuit/python.cpp:348-402 (raster_pairwise with network set,
                        circuitscape.cpp:474-743)

Pinned by Circuitscape 6.1's outputs for its network test INI files and the
synthetic ones (circuitscape_network.npz), by series / parallel resistances
and hand-solved circuits, and by the pseudo-inverse formula
R_ij = P_ii + P_jj - 2 P_ij (tests/test_circuitscape_network.py).
"""

from __future__ import annotations

import numpy as np

SHIFT = 10 * np.finfo(np.float64).eps
POLICIES = ("keepall", "rmvsrc", "rmvgnd", "rmvall")


class Graph:
    """A network's graph: per node its neighbours (ascending) and summed conductances, its degree,
    and the components (labels in order of their smallest node)."""

    def __init__(self, a, b, w, m):
        a = np.asarray(a, dtype=np.int64)
        b = np.asarray(b, dtype=np.int64)
        w = np.asarray(w, dtype=np.float64)
        if np.any((a < 0) | (b < 0) | (a >= m) | (b >= m)):
            raise ValueError("node index out of range")
        self.m = int(m)
        # duplicates of one (row, column) summed in input order
        A: dict = {}
        for r, c, x in zip(a.tolist(), b.tolist(), w.tolist()):
            A[(r, c)] = A[(r, c)] + x if (r, c) in A else x
        both: dict = {}
        for (r, c), x in A.items():
            if r == c:
                continue
            for key in ((r, c), (c, r)):
                both[key] = both[key] + x if key in both else x
        self.nbr = [[] for _ in range(self.m)]
        for (r, c), x in sorted(both.items(), key=lambda kv: (kv[0][1], kv[0][0])):
            if x != 0:
                self.nbr[c].append((r, x))
        self.degree = np.array([sum(x for _, x in self.nbr[v]) for v in range(self.m)])
        self.comp = np.full(self.m, -1, dtype=np.int64)
        self.ncomp = 0
        for u in range(self.m):
            if self.comp[u] >= 0:
                continue
            self.comp[u] = self.ncomp
            stack = [u]
            while stack:
                v = stack.pop()
                for t, _ in self.nbr[v]:
                    if self.comp[t] < 0:
                        self.comp[t] = self.ncomp
                        stack.append(t)
            self.ncomp += 1

    def members(self, k):
        return np.flatnonzero(self.comp == k)

    def matrix(self, nodes, finite):
        """The Laplacian rows of ``nodes`` (the unknowns), finite grounds and the shift on the diagonal."""
        where = {int(v): i for i, v in enumerate(nodes)}
        A = np.zeros((len(nodes), len(nodes)))
        for i, v in enumerate(nodes):
            A[i, i] = self.degree[v] + (finite[v] if finite is not None else 0.0) + SHIFT
            for t, x in self.nbr[v]:
                if t in where:
                    A[i, where[t]] -= x
        return A


def node_currents(g: Graph, members, v, finite, cur):
    """max(inflow, outflow) per node; branches below 1e-8 of the largest in their direction ignored."""
    with np.errstate(divide="ignore", invalid="ignore"):  # a largest branch of 0 compares as inf or nan
        _node_currents(g, members, v, finite, cur)


def _node_currents(g: Graph, members, v, finite, cur):
    maxin = maxout = -np.inf
    for c in members:
        for r, x in g.nbr[c]:
            b = abs(x) * (v[r] - v[c])
            if r < c:
                maxin = max(maxin, b)
            else:
                maxout = max(maxout, b)
    for c in members:
        inflow = outflow = 0.0
        for r, x in g.nbr[c]:
            b = abs(x) * (v[r] - v[c])
            if b > 0:
                if not abs(b / maxin) < 1e-8:
                    inflow += b
            elif b < 0:
                if not abs(b / maxout) < 1e-8:
                    outflow -= b
        if finite is not None:
            f = finite[c] * v[c]
            if f < 0:
                inflow -= f
            elif f > 0:
                outflow += f
        cur[c] = inflow if inflow > outflow else outflow


def branch_currents(g: Graph, members, v):
    """Every edge once (node c and its smaller neighbour r): (r, c, |current|), values below 1e-8 of
    the largest signed one set to 0."""
    out = []
    top = -np.inf
    for c in members:
        for r, x in g.nbr[c]:
            if r >= c:
                continue
            b = abs(x) * (v[r] - v[c])
            top = max(top, b)
            out.append([r, c, b])
    for row in out:
        with np.errstate(divide="ignore", invalid="ignore"):
            small = abs(row[2] / top) < 1e-8
        if small:
            row[2] = 0.0
        row[2] = abs(row[2])
    return out


def network_advanced(a, b, w, m, source, ground, policy):
    g = Graph(a, b, w, m)
    src = np.array(source, dtype=np.float64)
    gnd = np.array(ground, dtype=np.float64)
    finite = np.where(gnd < np.inf, gnd, 0.0)
    any_finite = bool(np.any(finite != 0))
    p = POLICIES.index(policy)
    both = (src != 0) & (gnd != 0)
    if p in (1, 3):
        src[both] = 0.0
    elif p == 2:
        gnd[both] = 0.0
    gnd[(gnd == np.inf) & (src > 0)] = 0.0
    volt = np.zeros(m)
    solved = False
    for k in range(g.ncomp):
        mem = g.members(k)
        if not (src[mem].sum() != 0 and gnd[mem].sum() != 0):
            continue
        solved = True
        nodes = mem[gnd[mem] != np.inf]
        if len(nodes):
            A = g.matrix(nodes, finite if any_finite else None)
            volt[nodes] = np.linalg.solve(A, src[nodes])
    cur = np.zeros(m)
    node_currents(g, np.arange(m), volt, finite if any_finite else None, cur)
    br = branch_currents(g, np.arange(m), volt)
    brr = np.array([[r, c, x] for r, c, x in br]).reshape(-1, 3)
    return {"voltage": volt, "current": cur,
            "branch": (brr[:, 0].astype(np.int32), brr[:, 1].astype(np.int32), brr[:, 2]),
            "solved": solved, "n_components": g.ncomp}


def network_pairwise(a, b, w, m, points, excluded, options, emit=None):
    g = Graph(a, b, w, m)
    pts = [int(v) if 0 <= v < m else -1 for v in np.asarray(points).tolist()]
    n = len(pts)
    ex = np.asarray(excluded, dtype=np.uint8).reshape(n, n)
    maps = bool(options.get("volt_maps") or options.get("cur_maps") or options.get("cum") or options.get("max"))
    R = np.full((n, n), -1.0)
    cum = np.zeros(m)
    cum_branch: dict = {}
    positions: dict = {}
    for i, v in enumerate(pts):
        if v >= 0:
            positions.setdefault(v, []).append(i)
    for k in range(g.ncomp):
        csub = []
        for v in pts:
            if v >= 0 and g.comp[v] == k and v not in csub:
                csub.append(v)
        if not csub:
            continue
        for v in csub:
            pos = positions[v]
            for x in range(len(pos)):
                for y in range(x + 1, len(pos)):
                    R[pos[x], pos[y]] = R[pos[y], pos[x]] = 0.0
        jobs = [(csub[s], csub[d]) for s in range(len(csub)) for d in range(s + 1, len(csub))
                if any(not ex[ci, cj] for ci in positions[csub[s]] for cj in positions[csub[d]])]
        if not jobs:
            continue
        mem = g.members(k)
        ref = mem[-1]  # the component's last node is grounded
        nodes = mem[mem != ref]
        where = {int(v): i for i, v in enumerate(nodes)}
        X = np.zeros((len(nodes), len(jobs)))
        if len(nodes):
            B = np.zeros((len(nodes), len(jobs)))
            for j, (s, d) in enumerate(jobs):
                if s in where:
                    B[where[s], j] = -1.0
                if d in where:
                    B[where[d], j] = 1.0
            X = np.linalg.solve(g.matrix(nodes, None), B)
        for j, (s, d) in enumerate(jobs):
            v = np.zeros(m)
            v[nodes] = X[:, j]
            vcol = v[mem] - v[s]
            resistance = (v[d] - v[s]) - 0.0
            if maps:
                vn = np.zeros(m)
                vn[mem] = vcol
                cn = np.zeros(m)
                node_currents(g, mem, vn, None, cn)
                br = branch_currents(g, mem, vn)
            for ci in positions[s]:
                for cj in positions[d]:
                    if ex[ci, cj]:
                        continue
                    R[ci, cj] = R[cj, ci] = resistance
                    if not maps:
                        continue
                    cum[mem] += cn[mem]
                    for r, c, x in br:
                        cum_branch[(r, c)] = cum_branch.get((r, c), 0.0) + x
                    if emit is not None:
                        brr = np.array(br).reshape(-1, 3)
                        emit(ci, cj, mem.astype(np.int32), vcol.copy() if options.get("volt_maps") else None,
                             cn[mem].copy(), (brr[:, 0].astype(np.int32), brr[:, 1].astype(np.int32), brr[:, 2]))
    np.fill_diagonal(R, 0.0)
    out = {"resistances": R, "n_components": g.ncomp}
    if maps:
        out["cum"] = cum
        keys = [(r, c) for c in range(m) for r, _ in g.nbr[c] if r < c]
        out["cum_branch"] = (np.array([r for r, _ in keys], dtype=np.int32),
                             np.array([c for _, c in keys], dtype=np.int32),
                             np.array([cum_branch.get(key, 0.0) for key in keys]))
    return out
