/*
 * FP64 CPU oracle for the STAQ-DIC hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This restatement is the checker the parity tests
 * compare the HIP library against, and the CPU baseline that bench.py times
 * beside the GPU.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product path (staq_dic_amd) never does.
 *
 * Array conventions follow the reference exactly (MATLAB, after the transpose on
 * load in func_quadtree/ReadImageQuadtree.m:78): every image is column-major with
 * dim 1 = x (H entries) and dim 2 = y (W entries); coordinates are 1-based.
 * Element (x, y) lives at a[(x-1) + (y-1)*H].  Gradient arrays are (H-6) x (W-6)
 * and element (i, j) belongs to pixel (i+3, j+3) (func/funImgGradient.m:53-78).
 */
#ifndef STAQ_ORACLE_H
#define STAQ_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    const double *f;     /* reference image, H x W                         */
    const double *g;     /* deformed image, H x W                          */
    const double *mask;  /* reference-image mask (0/1), H x W              */
    const double *dfdx;  /* d f / d x, (H-6) x (W-6), already masked       */
    const double *dfdy;  /* d f / d y, (H-6) x (W-6), already masked       */
    long H, W;
} orc_frame;

/* Status codes of the reference (funICGNQuadtree.m:422-442, LocalICGNQuadtree.m:92-97). */
enum { ORC_ST_NOTCONV = 101, ORC_ST_MASKED = 102, ORC_ST_FEATURELESS = 103, ORC_ST_ERROR = -1 };

/* ba_interp2.cpp:184-234 — Keys (a=-0.5) bicubic with clamp-to-edge.
 * F is M x N column-major; X is the dim-2 (column) coordinate, Y the dim-1 (row)
 * coordinate, both 1-based.  Writes nd values. */
void orc_ba_interp2(const double *F, long M, long N, const double *X, const double *Y,
                    long nd, double *out);

/* func/funImgGradient.m:53-78 — 7-tap central difference, crop 3, times mask. */
void orc_img_gradient(const double *img, const double *mask, long H, long W,
                      double *dfdx, double *dfdy);

/* bwselect(BW, c, r, conn) restated as a flood fill (c = column, r = row,
 * 1-based).  bw/out are nr x nc column-major bytes. */
void orc_bwselect(const unsigned char *bw, long nr, long nc, long c, long r, int conn,
                  unsigned char *out);

/* findpeak.m:1-68 with subpixel = 1.  f is nr x nc column-major.
 * Returns xpeak (column), ypeak (row), max_f. */
void orc_findpeak(const double *f, long nr, long nc, double *xpeak, double *ypeak,
                  double *max_f);

/* normxcorr2(C, D) restricted to the block funIntegerSearch.m:134 keeps:
 * shifts s1, s2 in [-1, 2r+1], output (2r+3) x (2r+3) column-major.
 * C is m1 x m2, D is (m1+2r) x (m2+2r). */
void orc_normxcorr2_full(const double *C, long m1, long m2, const double *D, long d1, long d2,
                         double *out);
void orc_normxcorr2_kept(const double *C, long m1, long m2, const double *D, long r1,
                         long r2, double *out);

/* funIntegerSearch.m:3-35 + :82-203 (whole field, tempNoOfInitPt = 0).
 * gridx/gridy are the two-element ROI ranges; on return *nx, *ny hold the grid
 * size.  Outputs are ny x nx column-major (rows = y) and must hold max_pts.
 * Returns 0, or -1 if the grid would exceed max_pts. */
int orc_integer_search(const double *f, const double *g, long H, long W, long r1, long r2,
                       const long gridx[2], const long gridy[2], long winsize,
                       long winstepsize, long max_pts, long *nx, long *ny, double *x0,
                       double *y0, double *u, double *v, double *ccmax);

/* funICGNQuadtree.m:1-451 (plain = 0) or funICGN.m:1-323 (plain = 1).
 * x0, y0 are the already-rounded node coordinates.  Writes P[6] =
 * [u_x v_x u_y v_y u v], H21 (the packed lower triangle) and returns the status
 * code the batch driver stores in ConvItPerEle. */
int orc_icgn6(const orc_frame *fr, const double U0[2], double x0, double y0, long winsize,
              double tol, int plain, double P[6], double H21[21]);

/* funICGN_Subpb1.m:1-474 (Gauss-Newton).  Returns the status code; writes U[2].
 * An exception of the reference is returned as -1. */
int orc_icgn2_al(const orc_frame *fr, double x0, double y0, long wx, long wy, double mu,
                 const double vdual[2], const double UOld[2], const double FOld[4],
                 double tol, double U[2]);

/* Batch drivers: one subset per OpenMP iteration, schedule(dynamic), like the
 * reference's parfor (LocalICGNQuadtree.m:108-139, Subpb1Quadtree.m:125-158).
 * coords is N x 2 column-major (coordinatesFEM); nodes are rounded here
 * (LocalICGNQuadtree.m:69) unless plain.  winsize per node (N entries).
 * Outputs: U (2N interleaved), F (4N), H (N x 21 column-major), status (N). */
void orc_local_icgn_batch(const orc_frame *fr, const double *U0, const double *coords,
                          const long *winsize, long N, double tol, int plain, int nthreads,
                          double *U, double *F, double *Hpar, int *status);
void orc_subpb1_batch(const orc_frame *fr, const double *UOld, const double *FOld,
                      const double *vdual, const double *coords, const long *wx,
                      const long *wy, long N, double mu, double tol, int plain, int nthreads,
                      double *U, int *status);

/* Total ZNSSD mismatch of a batch of affine subsets and its gradient.
 * C_i = sum_p ((f_p - fbar)/bf - (g(W(P_i;p)) - gbar)/bg)^2 with whole-window
 * statistics; grad (6N, may be NULL) is dC/dP_i.  Returns sum_i C_i; per-subset
 * values go to cost (may be NULL).  status_out (may be NULL) = 0 ok, 1 out of
 * image. */
double orc_znssd(const orc_frame *fr, const double *P, const double *coords,
                 const long *winsize, long N, double *cost, double *grad, int *status_out);

#ifdef __cplusplus
}
#endif
#endif
