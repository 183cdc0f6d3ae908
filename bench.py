"""Benchmark: converged IC-GN subsets/s of one batched local step on one GPU.

Headline workload (the large-deformation configuration, the largest that
fits one GPU): frames 39 -> 40 of a 4096 x 4096 16-bit speckle sequence
(192 000 Gaussian speckles, seed 1004) under a cumulative stretch of up to
20 % plus a 40 px Gaussian bump; 98 473 quadtree nodes (leaves 4-64 px).
Every node gets the one window the reference's quadtree drivers give it
(DICpara.winsize = 32, LocalICGNQuadtree.m:85), grown by the IC-GN step where
it meets a mask (none here).  A step is one 6-DOF IC-GN local step
(LocalICGNQuadtree) over every subset: one hipGraph holding the per-bin
launches and the FP64 re-solve, inputs already resident in HBM, tol 1e-3, at
most 100 iterations.  The value counts the converged subsets (status 1-100),
whose RMS error against the analytic field is the accuracy gate.

The JSON line also carries: a first pass of the same step (eager, default
forms, no ordering) beside the measured schedule; the same nodes with
windows tied to their leaf size (16-64 px) and the integer searches of the
pair (whole-field w 64 r 48, multigrid); the quadtree-with-hole pair (1024^2,
6016 nodes, reference and leaf window rules, the 2-DOF augmented-Lagrangian
step, the one-launch-per-subset form's summed device time, the integer
search); the plain method's uniform grid; the adjoint ZNSSD forward+backward
on the 2048^2 sinusoidal pair; the multi-instance batch (16 pairs of
512^2-2048^2 in one graph, per-pair latency percentiles); one frame pair
through the incremental driver's AL loop on the reference's own frames; the
roofline of the step's dominant launch (HBM, and its vector-ALU rate from the
committed counter summary); the CPU baseline (the FP64 oracle on every host
core, also on one core, and its integer search) and the host times of the
global step and the nodal strain.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "converged IC-GN subsets/s at ≤0.01 px RMS, batched vs plain; % HBM peak"
HBM_PEAK_GBS = 8000.0
# 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz, one lane operation per cycle
VALU_PEAK_TOPS = 1024 * 16 * 2.4e9 / 1e12


def algorithmic_bytes(wx, wy, iters):
    """B(w, k) = 4 [3 (w+1)^2 + k (w+4)^2] + 104 per subset (FP32 f, df/dx,
    df/dy read once; the bicubic support of g once per iteration; P in/out,
    status)."""
    nf = (wx + 1.0) * (wy + 1.0)
    ng = (wx + 4.0) * (wy + 4.0)
    return float(np.sum(4.0 * (3.0 * nf + iters * ng) + 104.0))


def _profiled(key, kernel, path="traffic_large_deformation.md"):
    """A per-launch figure of `kernel` from a committed counter summary under
    profiles/ (lines "<key> <kernel> <value>"), or None."""
    path = os.path.join(ROOT, "profiles", path)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + " ") and kernel in line:
            try:
                return float(line.split()[-1])
            except ValueError:
                return None
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_name(bin_desc, mode=6):
    f = bin_desc["bs6" if mode == 6 else "bs2"]
    if f in (1, 2, 4, 8, 16):
        return f"icgn_lean_kernel<{bin_desc['ppt']},{mode},{f}>"
    if f <= 0:
        return f"icgn_wave_kernel<{bin_desc['ppt']},{mode},{1 if f == 0 else -f}>"
    return f"icgn_kernel<{f},{mode}>"


def accuracy(U, st, truth, plain=False):
    """RMS (px) of (u, v) against the analytic field over converged subsets
    (status 1-100), and over the subsets the reference keeps after its
    bad-subset rule (LocalICGNQuadtree.m:151-169: status > max(mean + std, 6)
    is refilled)."""
    from staq_dic_amd.api import classify_local
    conv = (st >= 1) & (st <= 100)
    err = np.hypot(U[0::2] - truth[:, 4], U[1::2] - truth[:, 5])
    bad, _ = classify_local(st, plain=plain)
    kept = conv & ~bad
    rms = float(np.sqrt(np.mean(err[conv] ** 2))) if conv.any() else float("nan")
    rms_kept = float(np.sqrt(np.mean(err[kept] ** 2))) if kept.any() else float("nan")
    return conv, err, rms, rms_kept, kept


class Timer:
    """HIP events on the context's stream (the stream every launch of a step
    is enqueued on, or forks from and joins)."""

    def __init__(self, stream):
        import torch
        self.torch = torch
        self.stream = stream
        self.e0 = torch.cuda.Event(enable_timing=True)
        self.e1 = torch.cuda.Event(enable_timing=True)

    def eager_ms(self, b, p):
        """One eager step (every launch enqueued by the host), timed."""
        self.torch.cuda.synchronize()
        self.e0.record(self.stream)
        b.icgn6(p)
        self.e1.record(self.stream)
        self.torch.cuda.synchronize()
        return self.e0.elapsed_time(self.e1)

    def graph_ms(self, b, reps, warmup=2):
        for _ in range(warmup):
            b.graph_launch()
        self.torch.cuda.synchronize()
        self.e0.record(self.stream)
        for _ in range(reps):
            b.graph_launch()
        self.e1.record(self.stream)
        self.torch.cuda.synchronize()
        return self.e0.elapsed_time(self.e1) / reps


def schedule(b, p, mode, tm, tune=True):
    """Order each bin longest-first from a first pass, capture the step with
    the batch's default kernel forms, then (``tune``) with the autotuned
    forms, and keep whichever graph replays faster (the measured schedule)."""
    key = "bs6" if mode == 6 else "bs2"
    (b.icgn6 if mode == 6 else b.icgn2)(p)
    b.reorder()
    default_forms = [x[key] for x in b.bins()]
    b.graph_capture(p, mode)
    t_default = tm.graph_ms(b, 10)
    out = {"default_forms": default_forms, "default_graph_ms": t_default, "measured": "default"}
    if tune:
        b.autotune(p, mode)
        tuned = [x[key] for x in b.bins()]
        out["autotuned_forms"] = tuned
        b.graph_capture(p, mode)
        out["autotuned_graph_ms"] = tm.graph_ms(b, 10)
        if out["autotuned_graph_ms"] < t_default:
            out["measured"] = "autotuned"
        else:
            b.set_forms(default_forms, mode)
            b.graph_capture(p, mode)
    log(f"[bench] mode {mode} schedule: {out}")
    return out


def roofline(b, p, mode, st, win_x, win_y, steps, traffic=False):
    """Dominant launch of the step (the launch that carries the most
    algorithmic bytes), timed by events on the stream it runs on over
    ``steps`` eager steps."""
    bins = b.bins()
    kms = np.zeros(len(bins) + 1)
    for _ in range(steps):
        kms += b.step_timed(p, mode)
    kms /= steps
    bin_of = b.bin_of()
    conv = (st >= 1) & (st <= 100)
    iters = np.where(conv, st, 0).astype(float)
    bytes_bin = np.array([algorithmic_bytes(win_x[bin_of == k], win_y[bin_of == k],
                                            iters[bin_of == k]) for k in range(len(bins))])
    kdom = int(np.argmax(bytes_bin))
    kname = kernel_name(bins[kdom], mode)
    bytes_launch = float(bytes_bin[kdom])
    achieved = bytes_launch / (kms[kdom] / 1e3) / 1e9
    log(f"[bench] dominant launch {kname}: {kms[kdom]:.3f} ms, {bytes_launch / 1e6:.1f} MB, "
        f"{achieved:.1f} GB/s; launches {[round(float(x), 3) for x in kms]} ms")
    out = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS,
           "traffic": _profiled("traffic_bytes_per_launch", kname) if traffic else None,
           "kernel": kname, "subsets_per_launch": int((bin_of == kdom).sum()),
           "algorithmic_bytes_per_launch": bytes_launch, "launch_ms": float(kms[kdom]),
           "launch_ms_all": [float(x) for x in kms],
           "step_algorithmic_bytes": float(bytes_bin.sum())}
    if traffic:
        # the roof that binds: the vector ALU.  Wavefront VALU instructions
        # per launch from the committed SQ counter summary of this launch,
        # 64 lane operations each, against 1024 SIMDs x 16 lanes x 2.4 GHz
        # (no packed FP32)
        valu = _profiled("valu_insts_per_launch", kname, "sq_counters_large_deformation.md")
        if valu:
            lane_ops = valu * 64 / (kms[kdom] / 1e3)
            out["valu"] = {"bound": "valu", "achieved": lane_ops / 1e12, "peak": VALU_PEAK_TOPS,
                           "unit": "T lane-ops/s", "frac": lane_ops / 1e12 / VALU_PEAK_TOPS,
                           "valu_insts_per_launch": valu,
                           "valu_insts_per_subset": valu / max(int((bin_of == kdom).sum()), 1)}
    return out


def local_step(ctx, tm, coords, win, truth, steps, warmup, slot=None, tune=True, roof=True,
               traffic=False, plain=False, first_pass=False):
    """Build a batch, schedule it, replay the captured 6-DOF step ``steps``
    times; returns (result dict, batch).  ``plain``: the uniform-grid twin
    (LocalICGN / funICGN semantics)."""
    from staq_dic_amd import Batch, params
    from staq_dic_amd._lib import IN_U0
    b = Batch(ctx, coords, win, slot=slot)
    U0 = np.round(truth[:, 4:6]).reshape(-1)
    b.set_input(IN_U0, U0)
    p6 = params(tol=1e-3, plain=plain)
    first = None
    if first_pass:
        # a step of a fresh batch before any schedule: eager, default forms,
        # bins in the order set-up left them (no iteration counts known
        # yet); the second such step is timed, the first loads the kernels
        tm.eager_ms(b, p6)
        first = tm.eager_ms(b, p6)
    sched = schedule(b, p6, 6, tm, tune=tune)
    if first is not None:
        sched["first_pass_eager_ms"] = first
    ms = tm.graph_ms(b, steps, warmup)
    U, F, Hp, st = b.outputs()
    conv, err, rms, rms_kept, kept = accuracy(U, st, truth, plain)
    n_conv = int(conv.sum())
    out = {"subsets": int(len(coords)), "ms_per_step": ms, "converged": n_conv,
           "fp64_resolves": int(b.exact_count()),
           "converged_subsets_per_s": n_conv / (ms / 1e3), "rms_px": rms,
           "rms_px_kept": rms_kept, "kept": int(kept.sum()),
           "windows": {str(int(w)): int((win == w).sum()) for w in np.unique(win)},
           "schedule": sched}
    if roof:
        # the committed counter summary covers the headline workload only
        out["roofline"] = roofline(b, p6, 6, st, win.astype(float), win.astype(float), steps,
                                   traffic)
        out["roofline"]["step_achieved_gbs"] = out["roofline"]["step_algorithmic_bytes"] / (
            ms / 1e3) / 1e9
    log(f"[bench] icgn6 step {ms:.3f} ms, {n_conv}/{len(coords)} converged, rms {rms:.4f} px "
        f"(kept {rms_kept:.4f} px)")
    return out, b


def multi_pair_leg(ctx, tm, steps, warmup):
    """16 frame pairs resident at once (slots 0-15), all subsets in one batch
    and one graph; each pair's own step replayed alone gives its latency."""
    from staq_dic_amd import Batch, params, synth
    from staq_dic_amd._lib import IN_U0
    t0 = time.time()
    pairs = synth.config_multi()
    log(f"[bench] multi-pair workload ready: {sum(len(c) for _, c, _ in pairs)} subsets, "
        f"{time.time() - t0:.1f}s")
    for j, (fr, _, _) in enumerate(pairs):
        ctx.set_frame(fr.f, fr.g, fr.mask, slot=j)
    coords = np.concatenate([c for _, c, _ in pairs])
    win = np.concatenate([w for _, _, w in pairs])
    slot = np.concatenate([np.full(len(c), j, dtype=np.int32) for j, (_, c, _) in enumerate(pairs)])
    truth = np.concatenate([synth.truth_at(fr.field, c) for fr, c, _ in pairs])
    agg, b = local_step(ctx, tm, coords, win, truth, steps, warmup, slot=slot)
    b.close()
    p6 = params(tol=1e-3)
    lat = []
    for j, (fr, c, w) in enumerate(pairs):
        bj = Batch(ctx, c, w, slot=np.full(len(c), j, dtype=np.int32))
        bj.set_input(IN_U0, np.round(synth.truth_at(fr.field, c)[:, 4:6]).reshape(-1))
        bj.icgn6(p6)
        bj.reorder()
        bj.graph_capture(p6, 6)
        lat.append(tm.graph_ms(bj, max(3, steps // 4)))
        bj.close()
    lat = np.asarray(lat)
    agg.update({"workload": "16 pairs (seeds 5000-5015), sizes from {512,1024,1536,2048}^2, "
                            "random affine fields (|t| <= 10 px, |grad u| <= 0.05), quadtree "
                            "leaves 8-32 px, windows 16-64 px, 6-DOF local step, tol 1e-3",
                "pair_sizes": [int(fr.f.shape[0]) for fr, _, _ in pairs],
                "pair_subsets": [int(len(c)) for _, c, _ in pairs],
                "pair_latency_ms": [float(x) for x in lat],
                "pair_latency_ms_p50": float(np.percentile(lat, 50)),
                "pair_latency_ms_p99": float(np.percentile(lat, 99)),
                "pairs_in_sequence_ms": float(lat.sum())})
    log(f"[bench] multi-pair: one graph {agg['ms_per_step']:.3f} ms, per-pair p50 "
        f"{agg['pair_latency_ms_p50']:.3f} ms p99 {agg['pair_latency_ms_p99']:.3f} ms")
    return agg


def al_frame_leg(ctx):
    """Sections 4-6 of the incremental driver for the reference's frames
    000/001 (normalised over the ROI and masked as its driver does), started
    from the rounded stored answer; wall-clock seconds per kind of step."""
    from staq_dic_amd.aldic import ALDICQuadtreeFrame
    z = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden",
                             "circular_hole_000_001.npz"))
    gx, gy = z["gridx"], z["gridy"]

    def norm(img):
        t = img[gx[0] - 1:gx[1], gy[0] - 1:gy[1]].astype(np.float64)
        return (img.astype(np.float64) - t.mean()) / t.std(ddof=1)

    m0 = z["mask0"].astype(np.float64)
    f = norm(z["img0"]) * m0
    g = norm(z["img1"]) * z["mask1"].astype(np.float64)
    mesh = {"coordinatesFEM": z["coordinatesFEM"], "elementsFEM": z["elementsFEM"],
            "markCoordHoleEdge": z["markCoordHoleEdge"], "dirichlet": z["dirichlet"]}
    t = time.perf_counter()
    out = ALDICQuadtreeFrame(np.round(z["U1"]), mesh, {"ImgRefMask": m0}, f, g,
                             {"winsize": 32, "winstepsize": 16}, tol=1e-2, ctx=ctx)
    wall = time.perf_counter() - t
    U2 = out["USubpb2"]
    d = np.hypot(U2[0::2] - z["U1"][0::2], U2[1::2] - z["U1"][1::2])
    res = {"workload": "reference circular-hole frames 000/001, 1434-node quadtree mesh, "
                       "3 AL rounds, beta tuned over 3 values",
           "seconds": wall, "seconds_by_step": out["seconds"], "beta": out["beta"],
           "median_px_to_stored_answer": float(np.median(d))}
    log(f"[bench] AL frame {wall:.2f} s {out['seconds']}")
    return res


def large_leg(ctx, tm, steps, warmup, extras=True):
    """The headline: the large-deformation incremental pair (frames 39 -> 40
    of the 4096^2 16-bit sequence), 98 473 quadtree nodes with the
    reference's 32-px window, 6-DOF local step; with ``extras`` the same
    nodes with leaf-size windows and the pair's integer searches (w 64, r 48,
    and the multigrid search)."""
    from staq_dic_amd import synth
    from staq_dic_amd.api import integer_search_slot
    t0 = time.time()
    fr, coords, win_leaf = synth.config_large()
    truth = synth.truth_at(fr.field, coords)
    log(f"[bench] large-deformation workload ready: {len(coords)} subsets, "
        f"{time.time() - t0:.1f}s")
    H, W = fr.f.shape
    ctx.set_frame(fr.f, fr.g, fr.mask, slot=0)
    win = np.full(len(coords), 32, dtype=np.int64)
    head, b = local_step(ctx, tm, coords, win, truth, steps, warmup, traffic=True,
                         first_pass=True)
    b.close()
    log(f"[bench] headline step {head['ms_per_step']:.3f} ms "
        f"(first pass {head['schedule']['first_pass_eager_ms']:.3f} ms eager)")
    more = {}
    if extras:
        leaf, bl = local_step(ctx, tm, coords, win_leaf, truth, steps, warmup, tune=True,
                              roof=False)
        bl.close()
        leaf["workload"] = ("the same pair and nodes, windows tied to the leaf size "
                            "clamp(2 leaf, 16, 64)")
        more["leaf_window_rule"] = leaf
        # the search on the resident pair (its outputs come back to the host)
        x, y, u, v, cc = integer_search_slot(ctx, 0, 48, [1, H], [1, W], 64, 32)
        t = time.perf_counter()
        x, y, u, v, cc = integer_search_slot(ctx, 0, 48, [1, H], [1, W], 64, 32)
        dt = time.perf_counter() - t
        tr = synth.truth_at(fr.field, np.stack([x.ravel(order="F"), y.ravel(order="F")], axis=1))
        ok = np.isfinite(u.ravel(order="F"))
        du = np.abs(u.ravel(order="F")[ok] - tr[ok, 4])
        dv = np.abs(v.ravel(order="F")[ok] - tr[ok, 5])
        m, k = 65.0, 99.0
        search = {"points": int(x.size), "seconds": dt, "points_per_s": x.size / dt,
                  "fp64_tflops": x.size * k * k * m * m * 2 / dt / 1e12,
                  "fp64_peak_tflops": 78.6,
                  "within_1px": float(np.mean((du <= 1) & (dv <= 1)))}
        log(f"[bench] large integer search: {x.size} points in {dt * 1e3:.1f} ms")
        # the multigrid search of the incremental driver (InitFFTSearchMethod
        # = 0): whole-region FFT normxcorr2, the box levels, gridfit on the host
        from staq_dic_amd.mgsearch import mg_scatter, mg_search_slot
        t = time.perf_counter()
        tx, ty, tu, tv, phi, gx, gy = mg_scatter(ctx, 0, H, W, [1, H], [1, W], 64)
        t_lv = time.perf_counter() - t
        t = time.perf_counter()
        xg, yg, ug, vg, ccg = mg_search_slot(ctx, 0, H, W, [1, H], [1, W], 64, 32)
        t_all = time.perf_counter() - t
        trg = synth.truth_at(fr.field, np.stack([xg.ravel(), yg.ravel()], axis=1))
        okg = np.isfinite(ug.ravel())
        search["multigrid"] = {
            "boxes_last_level": int(tx.size), "seconds_levels": t_lv,
            "seconds_with_gridfit": t_all, "grid_points": int(ug.size),
            "within_1px": float(np.mean((np.abs(ug.ravel()[okg] - trg[okg, 4]) <= 1)
                                        & (np.abs(vg.ravel()[okg] - trg[okg, 5]) <= 1)))}
        log(f"[bench] multigrid search: levels {t_lv * 1e3:.1f} ms, with gridfit "
            f"{t_all * 1e3:.1f} ms")
        more["integer_search"] = search
    return head, more, (fr, coords, win, truth)


def hole_leg(ctx, tm, steps, warmup):
    """The quadtree-with-hole pair (1024^2, r = 150 hole, Kirsch field 8 px,
    6016 nodes): the 6-DOF step with the reference's 32-px windows and with
    leaf windows, the 2-DOF AL step, the one-launch-per-subset floor and the
    integer search; frame slot 1."""
    from staq_dic_amd import params, synth
    from staq_dic_amd._lib import IN_U0
    fr, coords, win_leaf = synth.config_hole()
    truth = synth.truth_at(fr.field, coords)
    win = np.full(len(coords), 32, dtype=np.int64)
    ctx.set_frame(fr.f, fr.g, fr.mask, slot=1)
    slot = np.full(len(coords), 1, dtype=np.int32)
    res, b = local_step(ctx, tm, coords, win, truth, steps, warmup, slot=slot, roof=True)
    res["workload"] = ("hole-quadtree-1024: 1024x1024 speckle pair (seed 1002), r=150 hole, "
                       "Kirsch field 8 px, quadtree leaves 8-32 px, every node 32 px "
                       "(DICpara.winsize), grown at the hole, 6-DOF local step, tol 1e-3")
    alt, balt = local_step(ctx, tm, coords, win_leaf, truth, steps, warmup, slot=slot,
                           tune=False, roof=False)
    balt.close()
    res["leaf_window_rule"] = alt
    # 2-DOF AL step on the 6-DOF result (identity global step, vdual = 0)
    U0 = np.round(truth[:, 4:6]).reshape(-1)
    b.set_input(IN_U0, U0)
    b.icgn6(params(tol=1e-3))
    b.chain()
    p2 = params(tol=1e-3, mu=1e-3)
    sched2 = schedule(b, p2, 2, tm)
    ms2 = tm.graph_ms(b, steps, warmup)
    _, _, _, st2 = b.outputs()
    c2 = int(((st2 >= 1) & (st2 <= 100)).sum())
    res["icgn2_al_step"] = {"ms": ms2, "converged": c2, "subsets_per_s": c2 / (ms2 / 1e3),
                            "schedule": sched2}
    log(f"[bench] hole icgn2 step {ms2:.3f} ms, {c2} converged")
    # one launch per subset (one workgroup each): summed device time
    b.set_input(IN_U0, U0)
    floor_ms = b.icgn6_per_subset(params(tol=1e-3), timed=True)
    _, _, _, stp = b.outputs()
    cp = int(((stp >= 1) & (stp <= 100)).sum())
    res["plain_form"] = {"kernel_ms_sum": floor_ms, "converged": cp,
                         "subsets_per_s": cp / (floor_ms / 1e3),
                         "batched_vs_plain": res["converged_subsets_per_s"] /
                                             (cp / (floor_ms / 1e3))}
    log(f"[bench] per-subset form floor {floor_ms:.1f} ms")
    b.close()
    # integer search on the same pair (w 32, step 16, r 10), frames resident
    from staq_dic_amd.api import integer_search_slot
    integer_search_slot(ctx, 1, 10, [1, 1024], [1, 1024], 32, 16)
    t = time.perf_counter()
    x, y, u, v, cc = integer_search_slot(ctx, 1, 10, [1, 1024], [1, 1024], 32, 16)
    dt = time.perf_counter() - t
    res["integer_search"] = {"points": int(x.size), "seconds": dt, "points_per_s": x.size / dt}
    return res, (fr, coords, win, truth)


def al_loop_hole_leg(ctx, al_steps=5):
    """One frame pair of the quadtree-with-hole workload (1024^2, Kirsch
    field) through the incremental driver's AL loop with 5 AL steps: the
    reference's quadtree mesh generated from the pair's mask (MeshSetUp,
    GenerateQuadtreeMesh), the rounded analytic field as the initial guess,
    local steps on a device-resident batch (captured graphs), thin-plate
    refit, beta tuning, global steps with the FE nodal strain on the host;
    seconds per kind of step and the RMS error of the result."""
    from staq_dic_amd import synth
    from staq_dic_amd.aldic import ALDICQuadtreeFrame, ResidentLocalSteps
    from staq_dic_amd.mesh import GenerateQuadtreeMesh, MeshSetUp, uniform_nodes
    fr, _, _ = synth.config_hole()
    H, W = fr.f.shape
    t = time.perf_counter()
    xn, yn = uniform_nodes(H, W, 32, 16)
    c0, e0 = MeshSetUp(xn, yn)
    coords, elems, _, mark = GenerateQuadtreeMesh(c0, e0, fr.mask, 16, 32)
    t_mesh = time.perf_counter() - t
    mesh = {"coordinatesFEM": coords, "elementsFEM": elems, "markCoordHoleEdge": mark,
            "dirichlet": mark}
    truth = synth.truth_at(fr.field, coords)
    U0 = np.round(truth[:, 4:6]).reshape(-1)
    t = time.perf_counter()
    res = ResidentLocalSteps(ctx, coords, 32, fr.f, fr.g, fr.mask)
    t_res = time.perf_counter() - t
    try:
        t = time.perf_counter()
        out = ALDICQuadtreeFrame(U0, mesh, {"ImgRefMask": fr.mask}, fr.f, fr.g,
                                 {"winsize": 32, "winstepsize": 16}, tol=1e-2, ctx=ctx,
                                 al_steps=al_steps, resident=res)
        wall = time.perf_counter() - t
    finally:
        res.close()
    U2 = out["USubpb2"]
    inside = fr.mask[np.round(coords[:, 0]).astype(int) - 1,
                     np.round(coords[:, 1]).astype(int) - 1] != 0
    err = np.hypot(U2[0::2] - truth[:, 4], U2[1::2] - truth[:, 5])[inside]
    r = {"workload": "hole-quadtree-1024 pair (seed 1002), reference quadtree mesh from its mask "
                     f"({len(coords)} nodes), {al_steps} AL steps, beta tuned over 3 values, "
                     "local steps on a resident batch (graphs), host FP64 global steps",
         "nodes": int(len(coords)), "seconds": wall, "seconds_by_step": out["seconds"],
         "seconds_mesh": t_mesh, "seconds_resident_setup": t_res, "beta": out["beta"],
         "ALSolveStep": out["ALSolveStep"], "bad_subsets_per_local_step": out["bad"],
         "rms_px_off_hole": float(np.sqrt(np.mean(err[np.isfinite(err)] ** 2)))}
    log(f"[bench] AL loop hole 1024: {wall:.2f} s {out['seconds']}, rms {r['rms_px_off_hole']:.4f} px")
    return r


def sequence_leg(ctx, n_frames=40, gn_steps=1, beta=1e-3):
    """The large-deformation sequence end to end: frames 0..n_frames of the
    4096^2 16-bit pattern under the cumulative stretch-and-bump field, each
    frame made on the device (inverse map by Newton's method, Keys bicubic of
    the pattern through staq_ba_interp2_dev, 16-bit quantisation, normalised
    over the image); for every increment k-1 -> k (reference frame updated
    each frame, as main_ALDIC_Quadtree_inc.m:64-73 does) the 6-DOF local step
    as one captured hipGraph replayed on the new pair, then gn_steps
    Gauss-Newton steps of the adjoint-fed global step (total ZNSSD plus the
    reference's F regulariser, beta) on the uniform 16-px FE mesh (MeshSetUp,
    64 516 nodes, windows 32).  Increment 1 starts from the rounded analytic
    increment, every later one from the previous increment's result, rounded.
    Per increment: local-step milliseconds, global-step seconds, RMS of both
    against the analytic increment over the converged subsets (and of the
    global step over every node, the non-converged ones included)."""
    import ctypes as C

    import torch
    from staq_dic_amd import Batch, params, synth
    from staq_dic_amd._lib import IN_U0
    from staq_dic_amd.adjoint_global import AdjointGlobalStep
    from staq_dic_amd.mesh import MeshSetUp, uniform_nodes
    H = W = 4096
    bits = 16
    t0 = time.perf_counter()
    f0 = synth.speckle_image(H, W, 192000, 1004, bits)
    t_pattern = time.perf_counter() - t0
    dev = torch.device("cuda")
    ext = torch.cuda.ExternalStream(ctx.stream)
    vp = lambda t: C.c_void_p(t.data_ptr())
    xn, yn = uniform_nodes(H, W, 32, 16)
    coords, el4 = MeshSetUp(xn, yn)
    N = len(coords)
    win = np.full(N, 32, dtype=np.int64)
    mesh = {"coordinatesFEM": coords,
            "elementsFEM": np.concatenate([el4, np.zeros((len(el4), 4), dtype=np.int64)], 1)}
    px = np.round(coords[:, 0]).astype(np.int64)
    py = np.round(coords[:, 1]).astype(np.int64)
    with torch.cuda.stream(ext):
        # images as x-fastest (W, H) tensors: the column-major H x W layout
        F0 = torch.from_numpy(np.ascontiguousarray(f0.T)).to(dev)
        del f0
        xg = torch.arange(1, H + 1, dtype=torch.float64, device=dev).expand(W, H).contiguous()
        yg = torch.arange(1, W + 1, dtype=torch.float64, device=dev)[:, None].expand(W, H)
        yg = yg.contiguous()
        ones = torch.ones((W, H), dtype=torch.float32, device=dev)

        def frame(k):
            s_k, amp, sig = 0.2 * k / 40, 40.0 * k / 40, 300.0
            xc = H / 2 + 0.5
            yc = W / 2 + 0.5
            # the field moves x only (v = 0), so y is its own pre-image and
            # Newton's method on X + u(X, y) = x is scalar
            X = xg.clone()
            for _ in range(30):
                gb = torch.exp(-((X - xc) ** 2 + (yg - yc) ** 2) / (2 * sig * sig))
                r = X + s_k * (X - xc) + amp * gb - xg
                dX = r / (1.0 + s_k - amp * gb * (X - xc) / (sig * sig))
                X = X - dX
                if float(dX.abs().max()) < 1e-10:
                    break
            Z = torch.empty_like(X)
            # ba_interp2(F, Xc, Yc): Xc the dim-2 (y) coordinate, Yc the dim-1 (x)
            ctx.check(ctx.L.staq_ba_interp2_dev(ctx.h, vp(F0), H, W, 1, vp(yg), vp(X), H * W,
                                                vp(Z), 0, C.c_void_p(ctx.stream)),
                      "staq_ba_interp2_dev")
            Z = torch.clamp(torch.round(Z), 0, (1 << bits) - 1)
            return ((Z - Z.mean()) / Z.std()).to(torch.float32)

        def set_pair(fa, fb):
            ctx.check(ctx.L.staq_set_frame_dev(ctx.h, 0, vp(fa), vp(fb), vp(ones), H, W,
                                               C.c_void_p(ctx.stream)), "staq_set_frame_dev")
            ctx.frames[0] = (H, W)

        def box_energy(fa):
            # sum (f - mean f)^2 over each node's window, from integral images (FP64)
            f64 = fa.to(torch.float64)
            I1 = torch.nn.functional.pad(f64.cumsum(0).cumsum(1), (1, 0, 1, 0))
            I2 = torch.nn.functional.pad((f64 * f64).cumsum(0).cumsum(1), (1, 0, 1, 0))
            x0 = torch.from_numpy(px - 16 - 1).to(dev)
            x1 = torch.from_numpy(px + 16).to(dev)
            y0 = torch.from_numpy(py - 16 - 1).to(dev)
            y1 = torch.from_numpy(py + 16).to(dev)
            box = lambda I: I[y1, x1] - I[y0, x1] - I[y1, x0] + I[y0, x0]
            s1, s2 = box(I1), box(I2)
            return (s2 - s1 * s1 / 33.0 ** 2).cpu().numpy()

        t0 = time.perf_counter()
        prev = frame(0)
        torch.cuda.synchronize()
        t_first_frame = time.perf_counter() - t0
        b = None
        p6 = params(tol=1e-3)
        rows = []
        U_prev = None
        for k in range(1, n_frames + 1):
            t_f = time.perf_counter()
            cur = frame(k)
            set_pair(prev, cur)
            torch.cuda.synchronize()
            t_f = time.perf_counter() - t_f
            fld = synth.IncrementalField(synth.stretch_bump_frame_field(k - 1, H, W),
                                         synth.stretch_bump_frame_field(k, H, W))
            truth = synth.truth_at(fld, coords)
            U0 = np.round(truth[:, 4:6]).reshape(-1) if U_prev is None else np.round(U_prev)
            if b is None:
                b = Batch(ctx, coords, win)
                b.set_input(IN_U0, U0)
                b.icgn6(p6)
                b.reorder()
                b.graph_capture(p6, 6)
            b.set_input(IN_U0, U0)
            ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            ev0.record(ext)
            b.graph_launch()
            ev1.record(ext)
            Ul, Fl, Hl, st = b.outputs()
            local_ms = ev0.elapsed_time(ev1)
            conv = (st >= 1) & (st <= 100)
            t_g = time.perf_counter()
            bf2 = box_energy(prev)
            D = np.zeros((N, 2, 2))
            D[:, 0, 0], D[:, 0, 1], D[:, 1, 1] = Hl[:, 18], Hl[:, 19], Hl[:, 20]
            D[:, 1, 0] = D[:, 0, 1]
            D = 2.0 * D / bf2[:, None, None]
            # subsets that did not converge: their neighbours' mean carries them
            Ug = np.where(np.repeat(conv, 2), Ul, np.nan).reshape(N, 2)
            Fg = np.where(np.repeat(conv, 4), Fl, np.nan).reshape(N, 4)
            bad = ~conv
            if bad.any():
                from scipy.spatial import cKDTree
                tree = cKDTree(coords[conv])
                for i in np.flatnonzero(bad):
                    nb = tree.query_ball_point(coords[i], 24.0)
                    if nb:
                        Ug[i] = Ug[conv][nb].mean(0)
                        Fg[i] = Fg[conv][nb].mean(0)
                    else:
                        Ug[i] = U0.reshape(N, 2)[i]
                        Fg[i] = 0.0
                D[bad] = np.median(D[conv], axis=0)
            Ua, costs = AdjointGlobalStep(mesh, ctx, win, Ug.ravel(), Fg.ravel(), beta, D,
                                          n_iter=gn_steps)
            t_g = time.perf_counter() - t_g
            err = lambda U: np.hypot(U[0::2] - truth[:, 4], U[1::2] - truth[:, 5])
            rows.append({"frame": k, "frame_s": t_f, "local_ms": local_ms,
                         "converged": int(conv.sum()), "global_s": t_g,
                         "rms_local_px": float(np.sqrt(np.mean(err(Ul)[conv] ** 2))),
                         "rms_global_px": float(np.sqrt(np.mean(err(Ua)[conv] ** 2))),
                         "rms_global_px_all_nodes": float(np.sqrt(np.mean(err(Ua) ** 2))),
                         "znssd_before_after": [costs[0], costs[-1]]})
            log(f"[bench] sequence frame {k}: local {local_ms:.2f} ms, {int(conv.sum())}/{N} "
                f"converged, rms local {rows[-1]['rms_local_px']:.4f} px, global "
                f"{rows[-1]['rms_global_px']:.4f} px ({rows[-1]['rms_global_px_all_nodes']:.4f} "
                f"px over all nodes, {t_g:.2f} s)")
            U_prev = Ua
            prev = cur
        b.close()
    loc = np.array([r["local_ms"] for r in rows])
    return {"workload": f"large-deformation-4096 sequence, frames 0-{n_frames} made on the "
                        "device, incremental pairs, uniform 16-px FE mesh (64 516 nodes, "
                        f"windows 32), 6-DOF local step as one replayed hipGraph, {gn_steps} "
                        f"Gauss-Newton step(s) of the adjoint-fed global step (beta {beta})",
            "frames": n_frames, "nodes": int(N), "seconds_pattern_cpu": t_pattern,
            "seconds_first_frame": t_first_frame,
            "local_ms_mean": float(loc.mean()), "local_ms_max": float(loc.max()),
            "local_subsets_per_s": float(N / (loc.mean() / 1e3)),
            "global_s_mean": float(np.mean([r["global_s"] for r in rows])),
            "frame_s_mean": float(np.mean([r["frame_s"] for r in rows])),
            "rms_local_px_max": float(max(r["rms_local_px"] for r in rows)),
            "rms_global_px_max": float(max(r["rms_global_px"] for r in rows)),
            "per_frame": rows}


def adjoint_bytes(w):
    """Algorithmic bytes of one forward + backward pass over a subset of
    window w: f ((w+1)^2 FP32) and the bicubic support of g ((w+4)^2 FP32)
    read by each pass, P in (48 B) each pass, the cost out (8 B), the
    upstream gradient in (8 B) and dC/dP out (48 B)."""
    return 8.0 * ((w + 1.0) ** 2 + (w + 4.0) ** 2) + 160.0


def adjoint_leg(ctx, tm, steps):
    import torch
    from staq_dic_amd import synth
    from staq_dic_amd.autograd import znssd
    fa, ca, wa = synth.config_adjoint()
    ta = synth.truth_at(fa.field, ca)
    N = len(ca)
    dev = torch.device("cuda", 0)
    stream = torch.cuda.ExternalStream(ctx.stream)
    ctx.set_frame(fa.f, fa.g, fa.mask, slot=0)
    wmax = int(wa.max())
    with torch.cuda.stream(stream):
        P = torch.tensor(ta, dtype=torch.float64, device=dev, requires_grad=True)
        cc_ = torch.tensor(ca, device=dev)
        ccm = cc_.t().contiguous()
        ww = torch.tensor(wa, dtype=torch.int32, device=dev)
        cost = torch.empty(N, dtype=torch.float64, device=dev)
        status = torch.empty(N, dtype=torch.int32, device=dev)
        gc = torch.ones(N, dtype=torch.float64, device=dev)
        gP = torch.empty((N, 6), dtype=torch.float64, device=dev)
        L, h, sp = ctx.L, ctx.h, ctx.stream

        def kernels():
            ctx.check(L.staq_znssd_fwd_dev(h, N, wmax, P.data_ptr(), ccm.data_ptr(), ww.data_ptr(),
                                           None, cost.data_ptr(), status.data_ptr(), sp))
            ctx.check(L.staq_znssd_bwd_dev(h, N, wmax, P.data_ptr(), ccm.data_ptr(), ww.data_ptr(),
                                           None, gc.data_ptr(), gP.data_ptr(), sp))

        for _ in range(3):
            kernels()
        torch.cuda.synchronize()
        tm.e0.record(stream)
        for _ in range(steps):
            kernels()
        tm.e1.record(stream)
        torch.cuda.synchronize()
        ms_k = tm.e0.elapsed_time(tm.e1) / steps
        for _ in range(2):
            c, _ = znssd(P, cc_, ww, ctx, wmax=wmax, coords_cm=ccm)
            c.sum().backward()
        torch.cuda.synchronize()
        tm.e0.record(stream)
        for _ in range(steps):
            P.grad = None
            c, _ = znssd(P, cc_, ww, ctx, wmax=wmax, coords_cm=ccm)
            c.sum().backward()
        tm.e1.record(stream)
        torch.cuda.synchronize()
    ms_a = tm.e0.elapsed_time(tm.e1) / steps
    byts = float(np.sum(adjoint_bytes(wa.astype(float))))
    ach = byts / (ms_k / 1e3) / 1e9
    log(f"[bench] adjoint fwd+bwd kernels {ms_k:.3f} ms, autograd {ms_a:.3f} ms for {N} subsets, "
        f"{ach:.0f} GB/s")
    return {"workload": "2048x2048 sinusoidal field (u = 4 sin(2 pi x/512), v = 3 cos(2 pi y/768)), "
                        "16-px grid, window 16",
            "subsets": N, "ms_fwd_bwd_kernels": ms_k, "ms_fwd_bwd_autograd": ms_a,
            "subsets_per_s": N / (ms_k / 1e3),
            "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": ach / HBM_PEAK_GBS, "algorithmic_bytes": byts,
                         "kernels": "znssd_wave_kernel<5,false> + znssd_wave_kernel<5,true>"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--no-sequence", action="store_true",
                    help="skip the 40-increment large-deformation sequence (about a minute)")
    ap.add_argument("--leg", choices=["adjoint", "al", "sequence"], default=None,
                    help="run one leg alone and print its JSON (a profiling aid)")
    ap.add_argument("--copy-calib", action="store_true",
                    help="first run a 256 MiB elementwise copy (read 256 MiB, write 256 MiB) "
                         "three times, the known-size kernel that scales the HBM counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this implementation runs on one GPU; sharding the subsets "
                                   "across GPUs is designed but not launched", "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    from staq_dic_amd import Context, synth

    if args.copy_calib:
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst

    if args.leg is not None:
        ctx = Context(0)
        if args.leg == "adjoint":
            out = adjoint_leg(ctx, Timer(torch.cuda.ExternalStream(ctx.stream)), args.steps)
        elif args.leg == "sequence":
            out = sequence_leg(ctx)
        else:
            out = al_loop_hole_leg(ctx)
        print(json.dumps(out))
        ctx.close()
        return

    ctx = Context(0)
    stream = torch.cuda.ExternalStream(ctx.stream)
    tm = Timer(stream)
    head, more, (fr, coords, win, truth) = large_leg(ctx, tm, args.steps, args.warmup,
                                                     extras=not args.no_extras)
    ms = head["ms_per_step"]
    N = len(coords)
    roof = head.pop("roofline")
    # value: converged subsets (status 1-100) per second; the accuracy gate is
    # their RMS error against the analytic field
    value = head["converged"] / (ms / 1e3)
    result = {
        "metric": METRIC, "value": value, "unit": "subsets/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "large-deformation-4096: frames 39 -> 40 of a 4096x4096 16-bit "
                               "speckle sequence (192 000 speckles, seed 1004), stretch up to 20 % "
                               "plus a 40 px Gaussian bump, 98 473 quadtree nodes (leaves 4-64 px), "
                               "every node 32 px (DICpara.winsize), 6-DOF IC-GN local step, "
                               "tol 1e-3, FP32 sums, FP64 6x6 solve",
                   "subsets": N, "parallelism": "1 GPU, one graph per local step"},
        "value_counts": "converged subsets (status 1-100) per second",
        "converged": head["converged"], "rms_px": head["rms_px"],
        "kept": head["kept"], "rms_px_kept": head["rms_px_kept"],
        "accuracy_gate_0p01px": {"all_converged": bool(head["rms_px"] <= 0.01),
                                 "kept_by_reference_rule": bool(head["rms_px_kept"] <= 0.01)},
        "fp64_resolves": head["fp64_resolves"],
        "schedule": head["schedule"],
        "schedule_note": "the measured graph's bin order (longest-first) comes from a prior "
                         "eager pass over the same inputs, and its kernel forms from the "
                         "autotuner on them; first_pass_eager_ms is a step before either",
        "roofline": roof,
    }
    result.update(more)
    ctx.close()

    if not args.no_extras:
        ctx = Context(0)
        tm = Timer(torch.cuda.ExternalStream(ctx.stream))
        hole, _ = hole_leg(ctx, tm, args.steps, args.warmup)
        result["hole_quadtree_1024"] = hole
        # the plain method on its uniform grid (512^2 affine pair, 900 windows of 32 px)
        fp, cp_, wp = synth.config_plain()
        ctx.set_frame(fp.f, fp.g, fp.mask, slot=3)
        plain_res, bp = local_step(ctx, tm, cp_, wp, synth.truth_at(fp.field, cp_), args.steps,
                                   args.warmup, slot=np.full(len(cp_), 3, dtype=np.int32),
                                   tune=False, roof=False, plain=True)
        bp.close()
        plain_res["workload"] = ("plain-512: 512x512 affine pair (seed 1001), uniform grid step 16, "
                                 "window 32, LocalICGN semantics, tol 1e-3")
        result["plain_method_512"] = plain_res

        # adjoint ZNSSD (forward + vector-Jacobian product through the bicubic
        # warp) on the 2048^2 sinusoidal-field pair with its 16-px node grid,
        # inputs resident: the two kernels called through the C ABI back to
        # back (device time), and the torch.autograd.Function around them
        result["adjoint_znssd"] = adjoint_leg(ctx, tm, args.steps)
        ctx.close()

        # multi-instance batch, in a fresh context
        ctx = Context(0)
        result["multi_pair"] = multi_pair_leg(ctx, Timer(torch.cuda.ExternalStream(ctx.stream)),
                                              args.steps, args.warmup)
        ctx.close()
        # one frame pair through the incremental driver's AL loop on the
        # reference's own circular-hole frames and 1434-node mesh: local steps
        # on the GPU, global step and FE nodal strain on the host
        ctx = Context(0)
        result["al_frame"] = al_frame_leg(ctx)
        result["al_loop_hole_1024"] = al_loop_hole_leg(ctx)
        ctx.close()
        if not args.no_sequence:
            # the 40 increments of the large-deformation sequence, local steps
            # replayed from one graph, adjoint-fed global steps on the host
            ctx = Context(0)
            seq = sequence_leg(ctx)
            ctx.close()
            seq["per_frame"] = [{k: (round(v, 6) if isinstance(v, float) else v)
                                 for k, v in r.items() if k != "znssd_before_after"}
                                for r in seq["per_frame"]]
            result["sequence_large_4096"] = seq

    if not args.no_cpu:
        from oracle import pyoracle
        of = pyoracle.Frame(fr.f, fr.g, fr.mask)
        U0 = np.round(truth[:, 4:6]).reshape(-1)
        nproc = os.cpu_count() or 1
        # the same 98 473-subset step, repeated for about cpu_seconds, on every
        # host core (the oracle's OpenMP loop, schedule(dynamic))
        reps, t_cpu, cconv = 0, 0.0, 0
        while t_cpu < args.cpu_seconds:
            t = time.perf_counter()
            _, _, _, so = pyoracle.local_icgn(of, U0, coords, win, tol=1e-3, nthreads=nproc)
            t_cpu += time.perf_counter() - t
            cconv += int(((so >= 1) & (so <= 100)).sum())
            reps += 1
        result["cpu_baseline"] = {"value": cconv / t_cpu, "unit": "subsets/s", "cores": nproc,
                                  "nproc": nproc, "kind": "port",
                                  "sample": f"{reps} x the full {N}-subset 6-DOF local step of "
                                            f"the headline workload (FP64 oracle, OpenMP "
                                            f"dynamic, {nproc} threads), {t_cpu:.1f} s"}
        log(f"[bench] cpu {cconv / t_cpu:.0f} subsets/s on {nproc} threads")
        # one core on every 16th subset of the same step
        sel = np.arange(0, N, 16)
        t = time.perf_counter()
        _, _, _, s1 = pyoracle.local_icgn(of, U0.reshape(-1, 2)[sel].reshape(-1), coords[sel],
                                          win[sel], tol=1e-3, nthreads=1)
        t1 = time.perf_counter() - t
        c1 = int(((s1 >= 1) & (s1 <= 100)).sum())
        # the integer search of the hole pair (w 32, step 16, r 10), all cores
        fh, _, _ = synth.config_hole()
        t = time.perf_counter()
        xs, _, _, _, _ = pyoracle.integer_search(fh.f, fh.g, 10, [1, 1024], [1, 1024], 32, 16)
        ts = time.perf_counter() - t
        result["cpu_baseline"].update({
            "single_core": {"value": c1 / t1, "unit": "subsets/s", "cores": 1,
                            "sample": f"every 16th subset ({sel.size}) of the same step, "
                                      f"{t1:.2f} s"},
            "integer_search": {"value": xs.size / ts, "unit": "points/s", "cores": nproc,
                               "sample": f"the {xs.size}-point search of the 1024^2 hole pair "
                                         f"(w 32, step 16, r 10), {ts:.2f} s"}})
        log(f"[bench] cpu one core {c1 / t1:.0f} subsets/s; search {xs.size / ts:.0f} points/s")

        # the host stages around the path (this package's own host code, FP64):
        # the global step and the FE nodal strain on the reference's 1434-node
        # quadtree mesh with its stored answer
        from staq_dic_amd.global_step import Subpb2Quadtree, funGlobalNodalStrainQuadtree
        z = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden",
                                 "circular_hole_000_001.npz"))
        mesh = {"coordinatesFEM": z["coordinatesFEM"], "elementsFEM": z["elementsFEM"],
                "dirichlet": z["dirichlet"]}
        t = time.perf_counter()
        Uh = Subpb2Quadtree(mesh, 2, 1e-3 * 16 ** 2, 1e-3, z["U1"], z["F1"], 0 * z["F1"],
                            0 * z["U1"], 0.0, 16.0, 0)
        t_s2 = time.perf_counter() - t
        t = time.perf_counter()
        funGlobalNodalStrainQuadtree(mesh, Uh, 2, 0)
        t_ns = time.perf_counter() - t
        result["host_stages"] = {"mesh": "reference circular-hole quadtree mesh, 1434 nodes",
                                 "subpb2_s": t_s2, "nodal_strain_s": t_ns}
        log(f"[bench] host Subpb2 {t_s2:.3f} s, nodal strain {t_ns:.3f} s")

    print(json.dumps(result))


if __name__ == "__main__":
    main()
