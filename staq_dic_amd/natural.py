"""Natural-neighbour (Sibson) interpolation with nearest-site extrapolation,
on the host (FP64): MATLAB's scatteredInterpolant(x, y, v, 'natural',
'nearest') as the reference calls it (func_quadtree/
funGlobalNodalStrainQuadtree.m:285-292, func_quadtree/
LocalICGNQuadtree.m:313-330, func_quadtree/Subpb1Quadtree.m:247-253).

Sibson's coordinates of a query q inside the convex hull of the sites: insert
q into the Voronoi diagram of the sites; the weight of site i is the area
that q's new cell takes from i's old cell, over the area of q's new cell.
Here that is computed without building either diagram whole:

1. the natural neighbours of q are the vertices of the Delaunay triangles
   whose circumcircle contains q (the Bowyer-Watson cavity, found by a walk
   from the triangle that contains q);
2. q's new cell is the intersection of the half-planes nearer to q than to
   each natural neighbour;
3. the natural neighbours proper are the candidates whose bisector with q
   carries an edge of that cell; the part of the cell that site i loses is
   the cell cut to i's old Voronoi cell, the half-planes nearer to i than to
   each of its Delaunay neighbours (co-circular ties only add half-planes
   that cut nothing).

The weights do not depend on how the triangulation breaks ties between
co-circular sites (the Voronoi diagram is unique): such a tie only adds
candidates whose share is empty.  A query on a site takes the site's value; a
query on the hull boundary takes the linear interpolation between the ends
of its hull edge (the limit of Sibson's coordinates there); a query outside
the hull takes the value of its nearest site, the lowest-numbered one where
several are equally near (MATLAB documents no tie rule).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
from scipy.spatial import Delaunay, cKDTree

__all__ = ["natural_nearest", "natural_plan", "sibson_weights"]

# ``_sibson`` below is the restatement in Python; the same steps run natively
# (staq_natural_interp) for the interior queries of ``natural_nearest``, and
# the tests hold the two to each other.


def _clip(poly, ax, ay, b):
    """Convex polygon (list of (x, y)) cut to the half-plane ax x + ay y <= b.
    Plain floats: the polygons here have a handful of vertices."""
    out = []
    k = len(poly)
    if k == 0:
        return out
    px, py = poly[-1]
    sp = ax * px + ay * py - b
    for cx, cy in poly:
        sc = ax * cx + ay * cy - b
        if sc <= 0.0:
            if sp > 0.0:
                t = sp / (sp - sc)
                out.append((px + t * (cx - px), py + t * (cy - py)))
            out.append((cx, cy))
        elif sp <= 0.0:
            t = sp / (sp - sc)
            out.append((px + t * (cx - px), py + t * (cy - py)))
        px, py, sp = cx, cy, sc
    return out


def _area(poly):
    if len(poly) < 3:
        return 0.0
    s = 0.0
    px, py = poly[-1]
    for cx, cy in poly:
        s += px * cy - py * cx
        px, py = cx, cy
    return 0.5 * abs(s)


class _Triangulation:
    def __init__(self, sites):
        self.p = np.asarray(sites, dtype=np.float64)
        self.tri = Delaunay(self.p)
        s = self.tri.simplices
        a, b, c = self.p[s[:, 0]], self.p[s[:, 1]], self.p[s[:, 2]]
        # circumcentres, relative to vertex a for accuracy
        bx, by = b[:, 0] - a[:, 0], b[:, 1] - a[:, 1]
        cx, cy = c[:, 0] - a[:, 0], c[:, 1] - a[:, 1]
        d = 2.0 * (bx * cy - by * cx)
        b2, c2 = bx * bx + by * by, cx * cx + cy * cy
        ux = (cy * b2 - by * c2) / d
        uy = (bx * c2 - cx * b2) / d
        self.cc = np.stack([a[:, 0] + ux, a[:, 1] + uy], axis=1)
        self.r2 = ux * ux + uy * uy
        span = np.ptp(self.p, axis=0).max()
        self.span = span if span > 0 else 1.0
        self.kd = cKDTree(self.p)
        self.pl = [(float(x), float(y)) for x, y in self.p]

    def cavity(self, q, s0):
        """Vertices of the triangles whose circumcircle holds q, by a walk
        over triangle neighbours from s0."""
        tol = 1e-10
        seen = {s0}
        stack = [s0]
        verts = set(self.tri.simplices[s0])
        nb = self.tri.neighbors
        while stack:
            t = stack.pop()
            for u in nb[t]:
                if u < 0 or u in seen:
                    continue
                seen.add(u)
                d = q - self.cc[u]
                if d @ d <= self.r2[u] * (1.0 + tol):
                    stack.append(u)
                    verts.update(self.tri.simplices[u])
        return np.array(sorted(verts))

    def on_hull(self, q, s):
        """(i, j, t) when q lies on a hull edge of triangle s, at
        q = (1-t) p_i + t p_j; a hull edge is one without a neighbour."""
        p = self.p
        eps = 1e-12 * self.span
        verts = self.tri.simplices[s]
        for k in range(3):
            if self.tri.neighbors[s][k] >= 0:
                continue
            i, j = verts[(k + 1) % 3], verts[(k + 2) % 3]
            e = p[j] - p[i]
            L2 = e @ e
            w = q - p[i]
            if abs(e[0] * w[1] - e[1] * w[0]) <= eps * np.sqrt(L2):
                t = (w @ e) / L2
                return i, j, min(max(t, 0.0), 1.0)
        return None


def _sibson(T, q, s0):
    """Indices and Sibson weights of the natural neighbours of interior q."""
    cand = T.cavity(q, s0)
    qx, qy = float(q[0]), float(q[1])
    pts = [(float(T.p[j, 0]), float(T.p[j, 1])) for j in cand]
    big = 1e4 * T.span
    cell = [(qx - big, qy - big), (qx + big, qy - big), (qx + big, qy + big), (qx - big, qy + big)]
    for (x, y) in pts:
        ax, ay = x - qx, y - qy
        cell = _clip(cell, ax, ay, ax * 0.5 * (x + qx) + ay * 0.5 * (y + qy))
    # the natural neighbours proper: the candidates whose bisector with q
    # carries an edge of the cell (two cell vertices on it); inside the cell
    # the nearest old site is always one of them
    cv = np.array(cell)
    P = T.p[cand]
    A = P - q
    S = cv @ A.T - (0.5 * ((P + q) * A).sum(axis=1))[None, :]
    tol = 1e-9 * T.span * np.sqrt((A * A).sum(axis=1))
    nn = np.flatnonzero((np.abs(S) <= tol[None, :]).sum(axis=0) >= 2)
    # site i's old cell is cut out by its Delaunay neighbours alone
    ptr, nbr = T.tri.vertex_neighbor_vertices
    areas = np.zeros(nn.size)
    for k, i in enumerate(nn):
        gi = cand[i]
        xi, yi = pts[i]
        poly = cell
        for gj in nbr[ptr[gi]:ptr[gi + 1]]:
            xj, yj = T.pl[gj]
            ax, ay = xj - xi, yj - yi
            poly = _clip(poly, ax, ay, ax * 0.5 * (xj + xi) + ay * 0.5 * (yj + yi))
            if len(poly) < 3:
                break
        areas[k] = _area(poly)
    tot = areas.sum()
    keep = areas > 0
    return cand[nn[keep]], areas[keep] / tot


def sibson_weights(sites, q):
    """(indices, weights) of Sibson's coordinates of one interior query q."""
    T = _Triangulation(sites)
    q = np.asarray(q, dtype=np.float64)
    s0 = int(T.tri.find_simplex(q[None, :])[0])
    if s0 < 0:
        raise ValueError("query outside the convex hull of the sites")
    return _sibson(T, q, s0)


def _native_interior(T, V, Q, simp):
    """Sibson interpolation of the interior queries Q (triangles simp) by
    staq_natural_interp in libstaqdic.so (the same steps as ``_sibson``)."""
    import ctypes as C

    from . import _lib
    L = _lib.load()
    tri = np.ascontiguousarray(T.tri.simplices, dtype=np.int32)
    nbr = np.ascontiguousarray(T.tri.neighbors, dtype=np.int32)
    ptr, adj = T.tri.vertex_neighbor_vertices
    ptr = np.ascontiguousarray(ptr, dtype=np.int64)
    adj = np.ascontiguousarray(adj, dtype=np.int32)
    P = np.ascontiguousarray(T.p, dtype=np.float64)
    Vc = np.ascontiguousarray(V, dtype=np.float64)
    Qc = np.ascontiguousarray(Q, dtype=np.float64)
    st = np.ascontiguousarray(simp, dtype=np.int32)
    out = np.empty((Qc.shape[0], Vc.shape[1]))
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
    rc = L.staq_natural_interp(P.shape[0], dp(P), tri.shape[0], ip(tri), ip(nbr),
                               ptr.ctypes.data_as(C.POINTER(C.c_long)), ip(adj), Vc.shape[1],
                               dp(Vc), Qc.shape[0], dp(Qc), ip(st), dp(out))
    if rc != 0:
        raise _lib.StaqError(f"staq_natural_interp failed ({rc})")
    return out


def _classify(T, Q):
    """Each query's class: ("site", i), ("out", i), ("edge", i, j, t) or
    ("in", triangle), by the rules of ``natural_nearest``."""
    simp = T.tri.find_simplex(Q)
    dist, near = T.kd.query(Q, k=1)
    out = []
    for k in range(Q.shape[0]):
        q = Q[k]
        if dist[k] <= 1e-12 * T.span:
            out.append(("site", int(near[k])))
            continue
        if simp[k] < 0:
            # outside the hull: the nearest site, the lowest index on a tie
            cand = T.kd.query_ball_point(q, dist[k] * (1 + 1e-12) + 1e-300)
            out.append(("out", int(min(cand)) if cand else int(near[k])))
            continue
        edge = T.on_hull(q, int(simp[k]))
        if edge is not None:
            out.append(("edge",) + tuple(edge))
            continue
        out.append(("in", int(simp[k])))
    return out


def natural_plan(sites, queries, cap=64):
    """The linear map of ``natural_nearest`` from site values to query values
    as a CSR matrix (queries x sites): one set of sites and queries, any
    number of value sets.  Each row holds its terms in the order
    ``natural_nearest`` adds them, so ``natural_plan(S, Q) @ V`` equals
    ``natural_nearest(S, V, Q)`` bit for bit."""
    import ctypes as C

    from . import _lib
    sites = np.asarray(sites, dtype=np.float64)
    Q = np.asarray(queries, dtype=np.float64)
    if sites.shape[0] < 3:
        raise ValueError("natural-neighbour interpolation needs at least three sites")
    T = _Triangulation(sites)
    cls = _classify(T, Q)
    nq = Q.shape[0]
    rows = [None] * nq
    interior = [k for k, c in enumerate(cls) if c[0] == "in"]
    if interior:
        L = _lib.load()
        ki = np.asarray(interior)
        tri = np.ascontiguousarray(T.tri.simplices, dtype=np.int32)
        nbr = np.ascontiguousarray(T.tri.neighbors, dtype=np.int32)
        ptr, adj = T.tri.vertex_neighbor_vertices
        ptr = np.ascontiguousarray(ptr, dtype=np.int64)
        adj = np.ascontiguousarray(adj, dtype=np.int32)
        P = np.ascontiguousarray(T.p)
        Qi = np.ascontiguousarray(Q[ki])
        st = np.ascontiguousarray([cls[k][1] for k in interior], dtype=np.int32)
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
        cnt = np.zeros(ki.size, dtype=np.int32)
        for _ in range(2):
            # a query inside a hole of the sites has the whole rim as
            # neighbours: a second call with room for the most of them
            idx = np.zeros((ki.size, cap), dtype=np.int32)
            w = np.zeros((ki.size, cap))
            rc = L.staq_natural_weights(P.shape[0], dp(P), tri.shape[0], ip(tri), ip(nbr),
                                        ptr.ctypes.data_as(C.POINTER(C.c_long)), ip(adj),
                                        ki.size, dp(Qi), ip(st), cap, ip(idx), dp(w), ip(cnt))
            if rc == 0 or int(cnt.max()) <= cap:
                break
            cap = int(cnt.max())
        if rc != 0:
            raise _lib.StaqError(f"staq_natural_weights failed ({rc})")
        for r, k in enumerate(interior):
            rows[k] = (idx[r, :cnt[r]], w[r, :cnt[r]])
    for k, c in enumerate(cls):
        if c[0] in ("site", "out"):
            rows[k] = (np.array([c[1]]), np.array([1.0]))
        elif c[0] == "edge":
            _, i, j, t = c
            rows[k] = (np.array([i, j]), np.array([1.0 - t, t]))
    indptr = np.zeros(nq + 1, dtype=np.int64)
    indptr[1:] = np.cumsum([r[0].size for r in rows])
    ind = np.concatenate([r[0] for r in rows]).astype(np.int64) if nq else np.zeros(0, np.int64)
    dat = np.concatenate([r[1] for r in rows]) if nq else np.zeros(0)
    return sp.csr_matrix((dat, ind, indptr), shape=(nq, sites.shape[0]))


def natural_nearest(sites, values, queries):
    """scatteredInterpolant(sites, values, 'natural', 'nearest') evaluated at
    queries: sites (n, 2), values (n,) or (n, m), queries (k, 2).  Interior
    queries go to the native routine in libstaqdic.so in one call."""
    sites = np.asarray(sites, dtype=np.float64)
    V = np.asarray(values, dtype=np.float64)
    squeeze = V.ndim == 1
    if squeeze:
        V = V[:, None]
    Q = np.asarray(queries, dtype=np.float64)
    out = np.empty((Q.shape[0], V.shape[1]))
    if sites.shape[0] < 3:
        raise ValueError("natural-neighbour interpolation needs at least three sites")
    T = _Triangulation(sites)
    interior, simp = [], []
    for k, c in enumerate(_classify(T, Q)):
        if c[0] in ("site", "out"):
            out[k] = V[c[1]]
        elif c[0] == "edge":
            _, i, j, t = c
            out[k] = (1.0 - t) * V[i] + t * V[j]
        else:
            interior.append(k)
            simp.append(c[1])
    if interior:
        ki = np.asarray(interior)
        out[ki] = _native_interior(T, V, Q[ki], np.asarray(simp))
    return out[:, 0] if squeeze else out

