"""ORACLE — test infrastructure only. numpy restatement of get_edge.m
(codes/GPU_implementation/seed/get_edge.m:16-121) and of the octahedral
sphere of sphere.m (codes/ACA_vs_CUR/seed/objects/sphere.m:13-53), used to
check the library's mesh and RWG tables.

Pins: get_edge.m's own invariants (one T+ and one T- per interior edge, the
free vertex is the triangle vertex not on the edge, trian = +e at T+ and -e at
T-) and Euler's formula, checked in tests/test_geometry.py.
"""
from __future__ import annotations

import numpy as np


def get_edge(vertex: np.ndarray, topol: np.ndarray) -> dict:
    """vertex 3 x Nv, topol 3 x Nt (1-based). Returns edges 4 x Ne, trian 3 x Nt,
    ln, cent, un, ds with get_edge.m semantics (no junctions)."""
    topol = np.asarray(topol, dtype=np.int64)
    v1, v2, v3 = topol[0], topol[1], topol[2]
    nt = v1.size
    t = np.arange(1, nt + 1)
    tt = np.concatenate([t, t, t])
    e = np.vstack([np.concatenate([v1, v2, v3]), np.concatenate([v2, v3, v1])])
    vv = np.concatenate([v3, v1, v2])
    ov = np.concatenate([np.full(nt, 3), np.full(nt, 1), np.full(nt, 2)])
    e1 = np.sort(e, axis=0)
    # sortrows(e1'): stable lexicographic on (first, second)
    ie2 = np.lexsort((e1[1], e1[0]), axis=0)
    e2 = e1[:, ie2].T
    same = np.all(e2[1:] == e2[:-1], axis=1)
    e4 = np.concatenate([[False], same])
    e5 = np.concatenate([e4[1:], [False]])
    tt1, tt2 = tt[ie2[e4]], tt[ie2[e5]]
    vv1, vv2 = vv[ie2[e4]], vv[ie2[e5]]
    ov1, ov2 = ov[ie2[e4]], ov[ie2[e5]]
    ne = tt1.size
    edges = np.vstack([tt1, tt2, vv1, vv2]).astype(np.float64)
    trian = np.zeros((3, nt))
    idx = np.arange(1, ne + 1)
    np.add.at(trian, (ov1 - 1, tt1 - 1), idx)
    np.add.at(trian, (ov2 - 1, tt2 - 1), -idx)
    a = vertex[:, e[0, ie2[e4]] - 1]
    b = vertex[:, e[1, ie2[e4]] - 1]
    ln = np.sqrt(np.sum((a - b) ** 2, axis=0))
    p1, p2, p3 = vertex[:, v1 - 1], vertex[:, v2 - 1], vertex[:, v3 - 1]
    cent = (p1 + p2 + p3) / 3
    c = np.cross((p3 - p1).T, (p2 - p1).T).T
    nc = np.sqrt(np.sum(c ** 2, axis=0))
    return {"edges": edges, "trian": trian, "ln": ln, "cent": cent, "un": c / nc, "ds": nc / 2}


def octasphere(radius: float, n_levels: int):
    """sphere.m: returns vertex 3 x Nv, topol 3 x Nt (1-based)."""
    V = [list(radius * np.array(p, dtype=float)) for p in
         ([0, 0, 1], [1, 0, 0], [0, 1, 0], [-1, 0, 0], [0, -1, 0], [0, 0, -1])]
    T = [[1, 3, 2], [1, 4, 3], [1, 5, 4], [1, 2, 5], [6, 2, 3], [6, 3, 4], [6, 4, 5], [6, 5, 2]]
    lookup = {tuple(v): i + 1 for i, v in enumerate(V)}

    def mid(a, b):
        r = [(V[a - 1][c] + V[b - 1][c]) / 2 for c in range(3)]
        nr = float(np.sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]))
        r = tuple(x * radius / nr for x in r)
        if r in lookup:
            return lookup[r]
        V.append(list(r))
        lookup[r] = len(V)
        return len(V)

    for _ in range(n_levels):
        nt = len(T)
        for t in range(nt):
            v1, v2, v3 = T[t]
            v12, v23, v31 = mid(v1, v2), mid(v2, v3), mid(v3, v1)
            T[t] = [v12, v23, v31]
            T.extend([[v2, v23, v12], [v23, v3, v31], [v31, v1, v12]])
    return np.array(V).T, np.array(T, dtype=float).T
