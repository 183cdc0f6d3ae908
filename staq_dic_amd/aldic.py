"""The augmented-Lagrangian loop of one frame pair (sections 4-6 of
main_ALDIC_Quadtree_inc.m:158-428): the local 6-DOF step on the GPU, the
global step and the FE nodal strain on the host, then up to two rounds of
the 2-DOF local step (GPU) and the global step.

Section 4 (:160-205): LocalICGNQuadtree from U0 (bad subsets refilled), then
  one thin-plate RBF through every finite node re-evaluated at all nodes
  (U and each F component separately).
Section 5 (:214-290): mu = 1e-3, zero duals; on the first frame pair beta is
  tuned over betaList = [1e-3, 1e-2, 1e-1] * winstepsize^2 * mu: for each, the
  global step and the nodal strain, Err1 = ||U1 - U2||, Err2 = ||F1 - F2||,
  each standardised (std with N-1), beta at the least sum, refined by the
  quadratic through its two neighbours in log10(beta) when both exist
  (MATLAB's fit 'poly2' through three points is that quadratic); the global
  step is redone with the chosen beta unless it is the last of the list.
  Later pairs take DICpara['beta'] (else 1e-3 * winstepsize^2 * mu).
  FSubpb2 is reset to FSubpb1 at the hole-edge nodes, USubpb2 and FSubpb2 at
  the nodes whose window touches a hole; the duals are F2 - F1 and U2 - U1.
Section 6 (:334-428): while ALSolveStep < 3: Subpb1Quadtree (windows
  DICpara['winsize'] everywhere) from (U2, F2) with the duals, FSubpb1 = F2,
  the global step and nodal strain, the same resets, the duals; the loop
  stops early when the RMS change of USubpb2 (or of USubpb1) between rounds
  is below 1e-2 and the pair is one where the driver measures it.

The nodal strain's component order is kept as the reference returns it (its
middle two entries are u_y, v_x where the local steps' F holds v_x, u_y), and
the driver mixes the two exactly as above.
"""
from __future__ import annotations

import time

import numpy as np

from .api import LocalICGNQuadtree, Subpb1Quadtree
from .global_step import Subpb2Quadtree, funGlobalNodalStrainQuadtree
from .refill import DENSE_MAX, thinplate_eval, thinplate_fill_local, thinplate_fit

__all__ = ["rbf_refit", "tune_beta", "ALDICQuadtreeFrame"]


def rbf_refit(U, F, coordinatesFEM):
    """main_ALDIC_Quadtree_inc.m:166-222: a thin-plate RBF through the nodes
    whose u is finite (notnanindex, :167), evaluated at every node, for u, v
    and each of the four F components; the F fits take the same centres
    (:192-222), so an F value at a node whose u is NaN is replaced by the fit
    and a NaN F at a node whose u is finite makes that fit NaN.  Past
    refill.DENSE_MAX centres the fit is the local-support form
    (refill.thinplate_fill_local) at the nodes whose u is NaN."""
    c = np.asarray(coordinatesFEM, dtype=np.float64)
    N = c.shape[0]
    U = np.asarray(U, dtype=np.float64).reshape(N, 2)
    F = np.asarray(F, dtype=np.float64).reshape(N, 4)
    ok = np.isfinite(U[:, 0])
    if ok.sum() > DENSE_MAX:
        # local support past the reference's mesh sizes: the nodes with a
        # finite u keep their values (the dense interpolant returns them), the
        # others take the thin-plate RBF through their nearest such nodes
        Uo, Fo = U.copy(), F.copy()
        miss = np.flatnonzero(~ok)
        Uo[miss] = thinplate_fill_local(c[ok], U[ok], c[miss])
        Fo[miss] = thinplate_fill_local(c[ok], F[ok], c[miss])
        return Uo.reshape(-1), Fo.reshape(-1)
    Uo = thinplate_eval(c[ok], thinplate_fit(c[ok], U[ok]), c)
    Fo = thinplate_eval(c[ok], thinplate_fit(c[ok], F[ok]), c)
    return Uo.reshape(-1), Fo.reshape(-1)


def tune_beta(errsum, betaList):
    """Best beta of main_ALDIC_Quadtree_inc.m:271-282: the least ErrSum, moved
    to the vertex of the quadratic through it and its two neighbours in
    log10(beta) when both neighbours exist."""
    i = int(np.argmin(errsum))
    if i < 1 or i > len(betaList) - 2:
        return float(betaList[i])
    x = np.log10(np.asarray(betaList[i - 1:i + 2], dtype=np.float64))
    p = np.polyfit(x, np.asarray(errsum[i - 1:i + 2], dtype=np.float64), 2)
    return float(10.0 ** (-p[1] / 2.0 / p[0]))


def _resets(U2, F2, U1, F1, hole_edge, hole_strain):
    U2 = U2.copy()
    F2 = F2.copy()
    for n in np.asarray(hole_edge, dtype=np.int64):
        F2[4 * n - 4:4 * n] = F1[4 * n - 4:4 * n]
    for n in np.asarray(hole_strain, dtype=np.int64):
        U2[2 * n - 2:2 * n] = U1[2 * n - 2:2 * n]
        F2[4 * n - 4:4 * n] = F1[4 * n - 4:4 * n]
    return U2, F2


class ResidentLocalSteps:
    """The local steps of one frame pair on a device-resident batch: the pair
    in frame slot 0 and the subset list (every node, window DICpara.winsize)
    stay in HBM, each kind of step is captured once into a hipGraph and
    replayed; only the step's inputs and outputs cross PCIe."""

    def __init__(self, ctx, coords, winsize, f, g, mask=None):
        from ._lib import Batch
        self.ctx = ctx
        ctx.set_frame(f, g, mask, slot=0)
        N = np.asarray(coords).shape[0]
        w = np.full(N, int(winsize), dtype=np.int64)
        self.b = Batch(ctx, np.asarray(coords, dtype=np.float64), w)
        self.graphs = {}

    def _run(self, mode, p):
        key = (mode, p.tol, p.mu)
        if self.graphs.get(mode) != key:
            (self.b.icgn6 if mode == 6 else self.b.icgn2)(p)  # a first eager step
            self.b.reorder()
            self.b.graph_capture(p, mode)
            self.graphs[mode] = key
        self.b.graph_launch()
        return self.b.outputs()

    def icgn6(self, U0, tol):
        from ._lib import IN_U0, params
        self.b.set_input(IN_U0, np.asarray(U0, dtype=np.float64))
        return self._run(6, params(tol=tol))

    def icgn2(self, UOld, FOld, vdual, mu, tol):
        from ._lib import IN_FOLD, IN_UOLD, IN_VDUAL, params
        self.b.set_input(IN_UOLD, np.asarray(UOld, dtype=np.float64))
        self.b.set_input(IN_FOLD, np.asarray(FOld, dtype=np.float64))
        self.b.set_input(IN_VDUAL, np.asarray(vdual, dtype=np.float64))
        U, _, _, st = self._run(2, params(tol=tol, mu=mu))
        return U, st

    def close(self):
        self.b.close()


def ALDICQuadtreeFrame(U0, DICmesh, Df, fNormalized, gNormalized, DICpara, tol=1e-2,
                       first_pair=True, measure_update=False, ctx=None, al_steps=3,
                       resident=None):
    """One frame pair through sections 4-6 of main_ALDIC_Quadtree_inc.m.

    DICmesh: coordinatesFEM, elementsFEM (E x 8), markCoordHoleEdge (also the
    global step's Dirichlet nodes).  DICpara: winsize, winstepsize, and beta
    for later pairs.  measure_update: whether the driver measures the change
    between rounds for this pair (:402), which allows the early stop.
    Returns a dict with USubpb2, FSubpb2 (the pair's result), USubpb1,
    FSubpb1, beta, ALSolveStep, the local steps' status codes and bad counts,
    and the seconds spent in each kind of step.  al_steps: the loop's bound
    (ALSolveStep < al_steps, 3 in the reference's drivers, :337).  resident:
    a ResidentLocalSteps of the pair, whose graphs then run the local steps."""
    coords = np.asarray(DICmesh["coordinatesFEM"], dtype=np.float64)
    N = coords.shape[0]
    mesh = {"coordinatesFEM": coords, "elementsFEM": DICmesh["elementsFEM"],
            "dirichlet": DICmesh.get("dirichlet", DICmesh.get("markCoordHoleEdge", [])),
            "neumann": DICmesh.get("neumann", np.zeros((0, 4)))}
    hole_edge = np.asarray(DICmesh.get("markCoordHoleEdge", []), dtype=np.int64)
    step = float(np.mean(DICpara["winstepsize"]))
    P = dict(DICpara)
    P.pop("winsize_List", None)
    times = {"local": 0.0, "global": 0.0, "refit": 0.0}

    # section 4
    t = time.perf_counter()
    raw = None if resident is None else resident.icgn6(U0, tol)
    U1, F1, Hp, _, conv1, nbad1, hole_strain = LocalICGNQuadtree(
        U0, coords, Df, fNormalized, gNormalized, P, "GaussNewton", tol, ctx=ctx, raw=raw)
    times["local"] += time.perf_counter() - t
    t = time.perf_counter()
    U1, F1 = rbf_refit(U1, F1, coords)
    times["refit"] += time.perf_counter() - t

    # section 5
    mu = 1e-3
    udual = 0 * F1
    vdual = 0 * U1
    order = 2
    alpha = 0.0
    betaList = np.array([1e-3, 1e-2, 1e-1]) * step ** 2 * mu
    t = time.perf_counter()
    if first_pair:
        err1 = np.zeros(3)
        err2 = np.zeros(3)
        for k, b in enumerate(betaList):
            U2 = Subpb2Quadtree(mesh, order, b, mu, U1, F1, udual, vdual, alpha, step, 0)
            F2, _, _ = funGlobalNodalStrainQuadtree(mesh, U2, order, 0)
            err1[k] = np.linalg.norm(U1 - U2)
            err2[k] = np.linalg.norm(F1 - F2)
        es = (err1 - err1.mean()) / err1.std(ddof=1) + (err2 - err2.mean()) / err2.std(ddof=1)
        beta = tune_beta(es, betaList)
    else:
        beta = float(P.get("beta", 1e-3 * step ** 2 * mu))
        U2 = F2 = None
    if U2 is None or abs(beta - betaList[-1]) > np.finfo(float).eps:
        U2 = Subpb2Quadtree(mesh, order, beta, mu, U1, F1, udual, vdual, alpha, step, 0)
        F2, _, _ = funGlobalNodalStrainQuadtree(mesh, U2, order, 0)
    times["global"] += time.perf_counter() - t
    U2, F2 = _resets(U2, F2, U1, F1, hole_edge, hole_strain)
    udual = F2 - F1
    vdual = U2 - U1

    # section 6
    status = [conv1]
    nbad = [nbad1]
    al = 1
    tol2 = 1e-2
    U2_prev, U1_prev = U2, U1
    P["winsize_List"] = np.full((N, 2), int(P["winsize"]), dtype=np.int64)
    while al < al_steps:
        al += 1
        t = time.perf_counter()
        raw = None if resident is None else resident.icgn2(U2, F2, vdual, mu, tol)
        U1, _, _, conv, nb = Subpb1Quadtree(U2, F2, udual, vdual, coords, Df, fNormalized,
                                            gNormalized, mu, beta, None, al, P, "GaussNewton",
                                            tol, ctx=ctx, raw=raw)
        times["local"] += time.perf_counter() - t
        F1 = F2
        status.append(conv)
        nbad.append(nb)
        t = time.perf_counter()
        U2 = Subpb2Quadtree(mesh, order, beta, mu, U1, F1, udual, vdual, alpha, step, 0)
        F2, _, _ = funGlobalNodalStrainQuadtree(mesh, U2, order, 0)
        times["global"] += time.perf_counter() - t
        U2, F2 = _resets(U2, F2, U1, F1, hole_edge, hole_strain)
        stop = False
        if measure_update:
            upd = np.linalg.norm(U2_prev - U2) / np.sqrt(U2.size)
            upd1 = np.linalg.norm(U1_prev - U1) / np.sqrt(U1.size)
            stop = upd < tol2 or upd1 < tol2
        U2_prev, U1_prev = U2, U1
        udual = F2 - F1
        vdual = U2 - U1
        if stop:
            break
    return {"USubpb2": U2, "FSubpb2": F2, "USubpb1": U1, "FSubpb1": F1, "beta": beta,
            "ALSolveStep": al, "status": status, "bad": nbad, "seconds": times,
            "markCoordHoleStrain": hole_strain}
