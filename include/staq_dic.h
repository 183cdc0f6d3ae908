/*
 * staq_dic.h — C ABI of the MI355X (gfx950) implementation of the STAQ-DIC
 * hot path: ZNCC integer search, 6-DOF and 2-DOF (augmented-Lagrangian)
 * IC-GN subset fitting, the bicubic gather, and the ZNSSD adjoint.
 *
 * Conventions (identical to the MATLAB reference, so MEX gateways pass mxArray
 * data straight through):
 *   - images are column-major, dim 1 = x (H entries), dim 2 = y (W entries),
 *     i.e. the arrays after the reference's transpose on load
 *     (func_quadtree/ReadImageQuadtree.m:78); coordinates are 1-based;
 *   - coordinatesFEM is N x 2 column-major (x then y);
 *   - U is 2N interleaved [u1 v1 u2 v2 ...], F is 4N [u_x v_x u_y v_y] per node,
 *     Hpar is N x 21 column-major (packed lower triangle of the 6x6 Hessian,
 *     funICGNQuadtree.m:448);
 *   - status is the reference's ConvItPerEle code: 1..100 iterations used,
 *     101 not converged / left the image / NaN, 102 >= 40 % masked,
 *     103 featureless, -1 an error the reference catches (LocalICGNQuadtree.m:92-97).
 *
 * Every call returns 0 on success and a negative STAQ_E* code on a fatal
 * error (bad arguments, out of memory, HIP failure); staq_last_error() then
 * describes it.  Per-subset failures are never fatal: they are status codes.
 * The caller owns every host buffer; outputs are preallocated by the caller.
 * A context is not re-entrant; distinct contexts may be used from distinct
 * threads.  Host-pointer calls are synchronous; *_dev and staq_batch_* calls
 * take device pointers / device-resident batches and are asynchronous on the
 * given stream (a hipStream_t passed as void*, NULL = the context stream).
 */
#ifndef STAQ_DIC_H
#define STAQ_DIC_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    STAQ_OK = 0,
    STAQ_EINVAL = -1,  /* bad argument                         */
    STAQ_ENOMEM = -2,  /* device or host allocation failed     */
    STAQ_EHIP = -3,    /* HIP runtime error                    */
    STAQ_ENOFRAME = -4 /* frame slot not set                   */
};

typedef struct staq_ctx staq_ctx;
typedef struct staq_batch staq_batch;

typedef struct staq_params {
    double tol;      /* IC-GN stopping threshold (LocalICGNQuadtree tol argument)        */
    double mu;       /* AL penalty of Subproblem 1 (Subpb1Quadtree mu)                   */
    int max_iter;    /* 100 in the reference (funICGNQuadtree.m:59)                      */
    int plain;       /* 1: uniform-grid twins funICGN / Subpb1 (no node rounding,
                        whole-window statistics, DfDx==0 mask rule)                      */
    int block_size;  /* kernel form for every bin: 0 = the batch's (tuned) schedule;
                        64/128/256/512 = workgroup kernel of that size; -1/-2/-4 = that
                        many wavefronts per subset; 1/2/4/8/16 = the lean kernel for
                        windows without masked pixels with that many wavefronts (a subset
                        it cannot finish goes to the fallback launch)                    */
    int persistent;  /* 1: persistent grid pulling subsets from a device work queue      */
    double guard;    /* guard band of the stopping tests: a subset whose |b|/|b_1| falls
                        within a factor (1 -/+ guard) of tol (|b|: 4 guard), whose warped
                        corner falls within 1e-3 px of the image test's bounds, or (6-DOF)
                        whose Hessian is indefinite or has Cholesky pivots spread by more
                        than 3e3, is solved again with FP64 arithmetic throughout (gather,
                        statistics, residual), so that FP32 sums never decide a stopping
                        or bounds test (funICGNQuadtree.m:154-157, 373-386;
                        funICGN_Subpb1.m:403-417); 0 = off                               */
} staq_params;

/* Defaults of the reference's drivers: tol 1e-2 (main_ALDIC_Quadtree_v2.m:212),
 * mu 1e-3 (:256), max_iter 100; guard 2e-3. */
void staq_params_default(staq_params *p);

/* ---- context ---------------------------------------------------------- */
int staq_ctx_create(staq_ctx **ctx, int device);
void staq_ctx_destroy(staq_ctx *ctx);
const char *staq_last_error(const staq_ctx *ctx);
/* hipStream_t of the context (as void*), for callers that sequence their own work. */
void *staq_ctx_stream(staq_ctx *ctx);

/* ---- frames: a reference/deformed pair resident in HBM as FP32 --------- */
/* f = fNormalized (already times its mask, main_ALDIC_Quadtree_v2.m:57-59),
 * g = gNormalized, mask = ImgRefMask (NULL = all ones); H x W column-major.
 * Computes Df (funImgGradient.m:53-78) on the device.  slot >= 0 selects the
 * frame pair (several pairs may be resident for batched launches). */
int staq_set_frame(staq_ctx *ctx, int slot, const double *f, const double *g, const double *mask,
                   long H, long W);
int staq_set_frame_f32(staq_ctx *ctx, int slot, const float *f, const float *g, const float *mask,
                       long H, long W);
/* Device-resident FP32 images (copied into the context's buffers). */
int staq_set_frame_dev(staq_ctx *ctx, int slot, const float *f, const float *g, const float *mask,
                       long H, long W, void *stream);
/* Copy the device gradient of a slot back: (H-6) x (W-6) each. */
int staq_get_gradient(staq_ctx *ctx, int slot, double *dfdx, double *dfdy);

/* ---- drop-ins of the reference's functions (host pointers) ------------- */

/* Z = ba_interp2(F, X, Y, 'cubic')  (ba_interp2.cpp:318-335).  F is M x N x P,
 * X (dim-2 coordinate) and Y (dim-1 coordinate) hold nd points; Z is nd x P. */
int staq_ba_interp2(staq_ctx *ctx, const double *F, long M, long N, long P, const double *X,
                    const double *Y, long nd, double *Z);
int staq_ba_interp2_f32(staq_ctx *ctx, const float *F, long M, long N, long P, const float *X,
                        const float *Y, long nd, float *Z);
/* Device pointers, FP32 or FP64 (single = 1 for FP32). */
int staq_ba_interp2_dev(staq_ctx *ctx, const void *F, long M, long N, long P, const void *X,
                        const void *Y, long nd, void *Z, int single, void *stream);

/* Df = funImgGradient(img, img, mask) (func/funImgGradient.m:53-78):
 * dfdx, dfdy are (H-6) x (W-6). */
int staq_img_gradient(staq_ctx *ctx, const double *img, const double *mask, long H, long W,
                      double *dfdx, double *dfdy);

/* [x,y,u,v,cc] = funIntegerSearch(f, g, [r1 r2], gridx, gridy, winsize, winstepsize, 0)
 * (func/funIntegerSearch.m:3-35, :82-203) on frame `slot` (f = its reference,
 * g = its deformed image).  Outputs are ny x nx column-major (rows = y) and
 * must hold max_pts entries; *nx, *ny receive the grid size. */
int staq_integer_search(staq_ctx *ctx, int slot, long r1, long r2, const long gridx[2],
                        const long gridy[2], long winsize, long winstepsize, long max_pts,
                        long *nx, long *ny, double *x0, double *y0, double *u, double *v,
                        double *ccmax);

/* The two correlations of the multigrid integer search funIntegerSearchMg
 * (func/funIntegerSearchMg.m:3-285, IntegerSearch.m:122-125 with
 * InitFFTSearchMethod = 0), on frame `slot`; boxes are [x0 x1 y0 y1], 1-based,
 * inclusive.  The level bookkeeping and the final gridfit are the caller's.
 *
 * staq_xcorr_full_peak: [xpeak, ypeak, max_f] = findpeak(normxcorr2(C, D), 1)
 * with C = f(box_f), D = g(box_g), the whole normxcorr2 map (:17-27); NaN
 * for a constant template (the reference's error). */
int staq_xcorr_full_peak(staq_ctx *ctx, int slot, const long box_f[4], const long box_g[4],
                         double out[3]);
/* staq_xcorr_boxes: for each of nbox boxes (4 longs each), C = f(box_f),
 * D = g(box_g +- pad), normxcorr2(C, D) cropped as :165 (2 pad + 3 rows,
 * 2 pad + 4 columns of placements), findpeak: out = [xpeak ypeak max_f] per box
 * (3 doubles), ok = 1, or ok = 0 and NaN where the reference's try fails
 * (D outside the image, template and g box of different sizes, constant
 * template). */
int staq_xcorr_boxes(staq_ctx *ctx, int slot, long nbox, const long *box_f, const long *box_g,
                     long pad, double *out, int *ok);

/* staq_xcorr_qfactors: the correlation quality factors of compute_qFactor
 * (funIntegerSearch.m:211-264) for nbox boxes (4 longs each, x0 x1 y0 y1,
 * 1-based, inclusive): C = f(box_f), D = g(box_g), the whole normxcorr2(C, D)
 * map ((n1 + d1 - 1) x (n2 + d2 - 1)), cc_min = map - min(map); q = [pce ppe]
 * per box: pce = max(cc_min)^2 / mean(cc_min.^2), ppe = 1 / the entropy of
 * cc_min's histogram in 30 equal bins over [0, max(cc_min)].  ok = 1, or
 * ok = 0 and NaN where the reference's try fails (a box outside the image, a
 * g box smaller than the template, a constant template).  The callers are the
 * whole-field search (:130-131, boxes ii:ii+w by ii-r:ii+w+r) and the last
 * level of the multigrid search (funIntegerSearchMg.m:164-173, g box +- 5). */
int staq_xcorr_qfactors(staq_ctx *ctx, int slot, long nbox, const long *box_f, const long *box_g,
                        double *q, int *ok);

/* staq_natural_interp (host only, no device work): Sibson natural-neighbour
 * interpolation at interior query points, the 'natural' method of MATLAB's
 * scatteredInterpolant as funGlobalNodalStrainQuadtree.m:285-292 and
 * LocalICGNQuadtree.m:313-330 call it.  sites n_sites x 2 and values
 * n_sites x m (row-major); tri / nbr: the Delaunay triangles (3 site indices
 * each) and their neighbours (-1 at the hull); vptr / vadj: each site's
 * Delaunay neighbours (CSR); qtri: the triangle holding each query (queries on
 * a site, on the hull or outside it are the caller's).  out: n_q x m. */
int staq_natural_interp(long n_sites, const double *sites, long n_tri, const int *tri,
                        const int *nbr, const long *vptr, const int *vadj, long m,
                        const double *values, long n_q, const double *queries, const int *qtri,
                        double *out);

/* staq_natural_weights (host only): the Sibson coordinates behind
 * staq_natural_interp, for callers that interpolate many value sets over one
 * set of sites and queries (the nodal strain of every global step of an AL
 * loop over one mesh, funGlobalNodalStrainQuadtree.m:285-292).  Same inputs
 * without the values; per query k, count[k] natural neighbours with positive
 * weight, their site indices idx[k * cap + i] and weights w[k * cap + i]
 * (i < min(count[k], cap)), in the order staq_natural_interp sums them.
 * Returns STAQ_EINVAL if any count exceeds cap (the counts are still set). */
int staq_natural_weights(long n_sites, const double *sites, long n_tri, const int *tri,
                         const int *nbr, const long *vptr, const int *vadj, long n_q,
                         const double *queries, const int *qtri, long cap, int *idx, double *w,
                         int *count);

/* Raw subset outputs of LocalICGNQuadtree (func_quadtree/LocalICGNQuadtree.m:63-147)
 * or, with params->plain, LocalICGN (func/LocalICGN.m:60-115), before the
 * bad-subset classification: U (2N), F (4N), Hpar (N x 21), status (N).
 * winsize is per node (N entries); slot may be NULL (all frame 0). */
int staq_local_icgn(staq_ctx *ctx, const staq_params *p, long N, const double *U0,
                    const double *coords, const long *winsize, const int *slot, double *U,
                    double *F, double *Hpar, int *status);

/* Raw outputs of Subpb1Quadtree (func_quadtree/Subpb1Quadtree.m:70-169) or,
 * with params->plain, Subpb1 (func/Subpb1.m): U (2N) and status (N; an error
 * the reference catches is -1 here).  udual is accepted for signature parity
 * and unused, as in funICGN_Subpb1. */
int staq_subpb1(staq_ctx *ctx, const staq_params *p, long N, const double *UOld,
                const double *FOld, const double *udual, const double *vdual,
                const double *coords, const long *wx, const long *wy, const int *slot,
                double *U, int *status);

/* No counterpart in the reference: staq_local_icgn for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Every pointer is a device pointer:
 * U0 (2N), coords (N x 2, column-major), winsize (N, int32, at most wmax),
 * outputs U (2N), F (4N), Hpar (N x 21), status (N), as staq_local_icgn
 * returns them; every subset on frame `slot`. The step's counters and list
 * of deferred subsets live in a fixed device array of the library (one per
 * GPU, about 4 MiB, kept until the process ends), so N is at most 1048576
 * and wmax at most 64 (STAQ_EINVAL beyond). Calls that overlap on one GPU
 * share that array and race: serialize them (one stream, or events).
 * Asynchronous on `stream` (NULL = the context stream). */
int staq_local_icgn_noalloc(staq_ctx *ctx, const staq_params *p, int slot, long N, int wmax,
                            const double *U0, const double *coords, const int *winsize,
                            double *U, double *F, double *Hpar, int *status, void *stream);

/* ---- ZNSSD mismatch and its adjoint (no reference counterpart) ---------- */
/* cost_i = sum_p (fhat - ghat(W(P_i)))^2 over the (w+1)^2 window of node i
 * (whole-window statistics); status 1 = window or warp outside the image.
 * P is 6N [u_x v_x u_y v_y u v] per subset. */
int staq_znssd_fwd(staq_ctx *ctx, long N, const double *P, const double *coords,
                   const long *winsize, const int *slot, double *cost, int *status);
/* Vector-Jacobian product: gradP (6N) = gcost_i * dcost_i/dP_i (gcost NULL =
 * ones); gradU (2N, may be NULL) = the translation part, i.e. the gradient
 * with respect to the nodal displacements the subsets are centred on. */
int staq_znssd_bwd(staq_ctx *ctx, long N, const double *P, const double *coords,
                   const long *winsize, const int *slot, const double *gcost, double *gradP,
                   double *gradU);
/* Device-pointer forms (winsize/slot int32; slot may be NULL). */
int staq_znssd_fwd_dev(staq_ctx *ctx, long N, int wmax, const double *P, const double *coords,
                       const int *winsize, const int *slot, double *cost, int *status,
                       void *stream);
int staq_znssd_bwd_dev(staq_ctx *ctx, long N, int wmax, const double *P, const double *coords,
                       const int *winsize, const int *slot, const double *gcost, double *gradP,
                       void *stream);

/* ---- device-resident batches (one launch per ADMM local step) ---------- */
/* A batch holds the subset list (coords N x 2, wx, wy (NULL = wx), slot (NULL
 * = 0)) in HBM together with its inputs and outputs, sorted into window-size
 * bins; launches are asynchronous on the context stream and may be captured
 * into a hipGraph. */
int staq_batch_create(staq_ctx *ctx, staq_batch **b, long N, const double *coords, const long *wx,
                      const long *wy, const int *slot);
void staq_batch_destroy(staq_batch *b);
enum { STAQ_IN_U0 = 0, STAQ_IN_UOLD = 1, STAQ_IN_FOLD = 2, STAQ_IN_VDUAL = 3 };
/* Upload an input vector (U0/UOld/vdual: 2N, FOld: 4N) from host memory. */
int staq_batch_set_input(staq_batch *b, int which, const double *host);
/* 6-DOF local step (LocalICGNQuadtree) over the whole batch. */
int staq_batch_icgn6(staq_batch *b, const staq_params *p);
/* 2-DOF AL local step (Subpb1Quadtree) over the whole batch. */
int staq_batch_icgn2(staq_batch *b, const staq_params *p);
/* One eager local step (mode 6 or 2) with every kernel launch bracketed by
 * events on the stream it runs on; synchronizes and writes the device time of
 * each bin's launch to kernel_ms[0 .. nbins-1] and of the fallback launch to
 * kernel_ms[nbins] (nbins from staq_batch_bins).  A measurement aid with no
 * counterpart in the reference. */
int staq_batch_step_timed(staq_batch *b, const staq_params *p, int mode, float *kernel_ms);
/* The one-workgroup-per-launch form: N launches of one subset each.  With
 * kernel_ms_sum non-NULL each launch is bracketed by events and the summed
 * device time (the overhead-free floor of this form) is returned. */
int staq_batch_icgn6_per_subset(staq_batch *b, const staq_params *p, double *kernel_ms_sum);
/* Copy outputs to host (any pointer may be NULL). */
int staq_batch_get_output(staq_batch *b, double *U, double *F, double *Hpar, int *status);
/* Scheduling from the previous step (results are unchanged): subsets the
 * lean kernel handed to the fallback launch move to the general kernel's bins
 * (the bins are then rebuilt with their default forms and a captured graph is
 * dropped), and each bin is ordered longest-first, predicted from the
 * iteration counts and window sizes. */
int staq_batch_reorder(staq_batch *b);
/* Feed the 6-DOF outputs (U, F) of this batch to its 2-DOF inputs (UOld, FOld)
 * on the device, vdual = 0 — the first AL step with an identity global step. */
int staq_batch_chain_outputs(staq_batch *b);
/* Time each candidate kernel form per window bin on the resident inputs
 * and keep the fastest (mode 6 or 2). */
int staq_batch_autotune(staq_batch *b, const staq_params *p, int mode);
/* Capture one local step (mode 6 or 2) into a hipGraph, then replay it. */
int staq_batch_graph_capture(staq_batch *b, const staq_params *p, int mode);
int staq_batch_graph_launch(staq_batch *b);
/* Number of subsets the last local step solved again in FP64 because a
 * stopping or bounds test fell inside the guard band (staq_params.guard).
 * A measurement aid with no counterpart in the reference. */
int staq_batch_exact_count(staq_batch *b, long *n);
/* Device pointers of the batch arrays (for callers that wrap them). */
typedef struct staq_batch_ptrs {
    double *coords, *U0, *UOld, *FOld, *vdual, *U, *F, *Hpar;
    int *status, *wx, *wy, *slot;
    long N;
    int wmax;
} staq_batch_ptrs;
int staq_batch_get_ptrs(staq_batch *b, staq_batch_ptrs *out);
/* Number of window bins and, for bin k, its window size cap, item count,
 * kernel form chosen for mode 6 / mode 2 (as staq_batch_set_form) and pixels
 * per lane of the wavefront kernel of the mode-6 form (0 = workgroup kernel). */
int staq_batch_bins(staq_batch *b, int *nbins, int *wcap, long *count, int *bs6, int *bs2,
                    int *ppt, int maxbins);

/* Set the kernel form of bin `bin` for mode 6 or 2 (0 = one wavefront per
 * subset, -1/-2/-4 = that many wavefronts, 1/2/4/8/16 = the lean kernel for
 * windows without masked pixels with that many wavefronts, 64..512 = workgroup
 * kernel), as staq_batch_autotune would; drops a captured graph. */
int staq_batch_set_form(staq_batch *b, int bin, int mode, int form);
/* Bin index of every subset (N entries, in the caller's subset order).  Bins
 * group similar window sizes; windows that touch the mask of their frame at
 * batch creation form bins of their own. */
int staq_batch_bin_of(staq_batch *b, int *bin);

#ifdef __cplusplus
}
#endif
#endif
