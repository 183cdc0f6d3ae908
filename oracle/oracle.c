/*
 * FP64 CPU oracle for the STAQ-DIC hot path — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's MATLAB functions, written from the
 * MATLAB sources as a specification.  It is the checker the parity tests use
 * and the CPU baseline bench.py times; the product path never loads it.
 * Each function cites the reference lines it follows.
 *
 * Parity pins (what checks each function against something other than itself)
 * are listed in tests/test_oracle.py: closed forms for the bicubic kernel and
 * findpeak's least-squares fit, brute-force ZNCC, known-answer warps for IC-GN,
 * forced status codes, and central finite differences for the ZNSSD gradient.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define IDX(a, H, x, y) ((a)[((x) - 1) + ((long)(y) - 1) * (H)])

/* ------------------------------------------------------------------------- */
/* Bicubic gather, ba_interp2.cpp:60-130 (clamped tap indices) and :184-234.  */
/* ------------------------------------------------------------------------- */

static inline long clampl(long v, long lo, long hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* Keys cubic-convolution weights with a = -0.5 for fractional offset t. */
static inline void keys_weights(double t, double w[4])
{
    const double t2 = t * t, t3 = t2 * t;
    w[0] = 0.5 * (-t + 2.0 * t2 - t3);
    w[1] = 0.5 * (2.0 - 5.0 * t2 + 3.0 * t3);
    w[2] = 0.5 * (t + 4.0 * t2 - 3.0 * t3);
    w[3] = 0.5 * (-t2 + t3);
}

/* d/dt of keys_weights. */
static inline void keys_dweights(double t, double w[4])
{
    const double t2 = t * t;
    w[0] = 0.5 * (-1.0 + 4.0 * t - 3.0 * t2);
    w[1] = 0.5 * (-10.0 * t + 9.0 * t2);
    w[2] = 0.5 * (1.0 + 8.0 * t - 9.0 * t2);
    w[3] = 0.5 * (-2.0 * t + 3.0 * t2);
}

/* One bicubic sample of F (M rows x N cols, column-major).  X = column
 * coordinate, Y = row coordinate, 1-based.  Taps are rows floor(Y)-1..floor(Y)+2
 * and columns floor(X)-1..floor(X)+2 (1-based), clamped to the array.  Inner sum
 * runs over columns (weights of X), outer over rows (weights of Y), as the
 * reference does (ba_interp2.cpp:226-230).  Optionally returns the partial
 * derivatives along X and Y. */
static double bicubic_sample(const double *F, long M, long N, double X, double Y, double *dX,
                             double *dY)
{
    if (isnan(X) || isnan(Y)) {
        if (dX) *dX = NAN;
        if (dY) *dY = NAN;
        return NAN;
    }
    const double xf = floor(X), yf = floor(Y);
    const double tx = X - xf, ty = Y - yf;
    double wx[4], wy[4], dwx[4], dwy[4];
    keys_weights(tx, wx);
    keys_weights(ty, wy);
    if (dX || dY) {
        keys_dweights(tx, dwx);
        keys_dweights(ty, dwy);
    }
    /* 0-based index of the floor pixel, then taps -1..+2 clamped. */
    const long cx = (long)xf - 1, cy = (long)yf - 1;
    double acc = 0.0, ax = 0.0, ay = 0.0;
    for (int j = 0; j < 4; ++j) {
        const long row = clampl(cy - 1 + j, 0, M - 1);
        double s = 0.0, sx = 0.0;
        for (int i = 0; i < 4; ++i) {
            const long col = clampl(cx - 1 + i, 0, N - 1);
            const double v = F[row + col * M];
            s += wx[i] * v;
            if (dX || dY) sx += dwx[i] * v;
        }
        acc += wy[j] * s;
        if (dX) ax += wy[j] * sx;
        if (dY) ay += dwy[j] * s;
    }
    if (dX) *dX = ax;
    if (dY) *dY = ay;
    return acc;
}

void orc_ba_interp2(const double *F, long M, long N, const double *X, const double *Y, long nd,
                    double *out)
{
    for (long i = 0; i < nd; ++i) out[i] = bicubic_sample(F, M, N, X[i], Y[i], NULL, NULL);
}

/* ------------------------------------------------------------------------- */
/* Image gradient, func/funImgGradient.m:53-78.                               */
/* imfilter with [-1/60 3/20 -3/4 0 3/4 -3/20 1/60]' along dim 1 (and its     */
/* transpose along dim 2) is a correlation: out(i) = sum_k h(k) I(i+k-4).     */
/* The crop (4:end-3) removes every zero-padded sample, so no padding occurs. */
/* ------------------------------------------------------------------------- */

void orc_img_gradient(const double *img, const double *mask, long H, long W, double *dfdx,
                      double *dfdy)
{
    static const double h[7] = {-1.0 / 60, 3.0 / 20, -3.0 / 4, 0.0, 3.0 / 4, -3.0 / 20, 1.0 / 60};
    const long Hd = H - 6;
    for (long y = 4; y <= W - 3; ++y)
        for (long x = 4; x <= H - 3; ++x) {
            double sx = 0.0, sy = 0.0;
            for (int k = 0; k < 7; ++k) {
                sx += h[k] * IDX(img, H, x + k - 3, y);
                sy += h[k] * IDX(img, H, x, y + k - 3);
            }
            const double m = IDX(mask, H, x, y);
            dfdx[(x - 4) + (y - 4) * Hd] = sx * m;
            dfdy[(x - 4) + (y - 4) * Hd] = sy * m;
        }
}

/* ------------------------------------------------------------------------- */
/* bwselect: the connected component of BW containing pixel (r, c).  A seed    */
/* on background, or outside the array, selects nothing.                      */
/* ------------------------------------------------------------------------- */

void orc_bwselect(const unsigned char *bw, long nr, long nc, long c, long r, int conn,
                  unsigned char *out)
{
    memset(out, 0, (size_t)(nr * nc));
    if (r < 1 || r > nr || c < 1 || c > nc) return;
    const long s = (r - 1) + (c - 1) * nr;
    if (!bw[s]) return;
    long *stack = (long *)malloc(sizeof(long) * (size_t)(nr * nc));
    long top = 0;
    stack[top++] = s;
    out[s] = 1;
    while (top > 0) {
        const long p = stack[--top];
        const long pr = p % nr, pc = p / nr;
        for (int dc = -1; dc <= 1; ++dc)
            for (int dr = -1; dr <= 1; ++dr) {
                if (!dr && !dc) continue;
                if (conn == 4 && dr && dc) continue;
                const long qr = pr + dr, qc = pc + dc;
                if (qr < 0 || qr >= nr || qc < 0 || qc >= nc) continue;
                const long q = qr + qc * nr;
                if (bw[q] && !out[q]) {
                    out[q] = 1;
                    stack[top++] = q;
                }
            }
    }
    free(stack);
}

/* ------------------------------------------------------------------------- */
/* findpeak.m:1-68 (subpixel = 1).                                           */
/* ------------------------------------------------------------------------- */

void orc_findpeak(const double *f, long nr, long nc, double *xpeak, double *ypeak, double *max_f)
{
    /* max(f(:)): first maximal element in column-major order; NaN ignored. */
    long imax = -1;
    double m = 0.0;
    for (long i = 0; i < nr * nc; ++i) {
        if (isnan(f[i])) continue;
        if (imax < 0 || f[i] > m) {
            m = f[i];
            imax = i;
        }
    }
    if (imax < 0) { /* all NaN: max returns NaN at index 1 */
        *max_f = NAN;
        *xpeak = 1;
        *ypeak = 1;
        return;
    }
    const long yp = imax % nr + 1, xp = imax / nr + 1; /* row, column (1-based) */
    *max_f = m;
    *xpeak = (double)xp;
    *ypeak = (double)yp;
    if (xp == 1 || xp == nc || yp == 1 || yp == nr) return; /* peak on edge */

    /* Least-squares fit u = A1 + A2 x + A3 y + A4 xy + A5 x^2 + A6 y^2 on the
     * 3x3 block (x = column offset, y = row offset).  On this grid the columns
     * x, y, xy are mutually orthogonal and orthogonal to 1, x^2, y^2, so the
     * normal equations split: A2 = sum(xu)/6, A3 = sum(yu)/6, A4 = sum(xyu)/4,
     * and [A1 A5 A6] solve [[9 6 6];[6 6 4];[6 4 6]] a = [su; sxxu; syyu],
     * whose inverse is [[5/9 -1/3 -1/3];[-1/3 1/2 0];[-1/3 0 1/2]]. */
    double su = 0, sxu = 0, syu = 0, sxyu = 0, sxxu = 0, syyu = 0;
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy) {
            const double u = f[(yp - 1 + dy) + (xp - 1 + dx) * nr];
            su += u;
            sxu += dx * u;
            syu += dy * u;
            sxyu += dx * dy * u;
            sxxu += dx * dx * u;
            syyu += dy * dy * u;
        }
    const double A1 = (5.0 * su - 3.0 * sxxu - 3.0 * syyu) / 9.0;
    const double A2 = sxu / 6.0, A3 = syu / 6.0, A4 = sxyu / 4.0;
    const double A5 = -su / 3.0 + sxxu / 2.0;
    const double A6 = -su / 3.0 + syyu / 2.0;
    const double den = A4 * A4 - 4.0 * A5 * A6;
    double xo = (-A3 * A4 + 2.0 * A6 * A2) / den;
    double yo = -1.0 / den * (A4 * A2 - 2.0 * A5 * A3);
    if (fabs(xo) > 1 || fabs(yo) > 1) return;
    xo = round(1000.0 * xo) / 1000.0; /* MATLAB round: half away from zero */
    yo = round(1000.0 * yo) / 1000.0;
    *xpeak = xp + xo;
    *ypeak = yp + yo;
    *max_f = A1 + A2 * xo + A3 * yo + A4 * xo * yo + A5 * xo * xo + A6 * yo * yo;
}

/* ------------------------------------------------------------------------- */
/* normxcorr2 on the kept block (funIntegerSearch.m:127,134).                 */
/* Published definition (Lewis 1995, MATLAB IPT documentation): with D zero-  */
/* padded, for each template placement                                        */
/*   num   = sum(T .* Dw) - sum(Dw) * sum(T) / mn                             */
/*   den   = sqrt(sum((T - mean T)^2)) * sqrt(max(sum(Dw.^2) - sum(Dw)^2/mn, 0))*/
/*   C     = num / den where den > sqrt(eps(max den)), else 0.                 */
/* The threshold max is taken over the kept block (the full map's maximum is  */
/* attained there for any zero-mean image; parity unpinned otherwise).        */
/* ------------------------------------------------------------------------- */

void orc_normxcorr2_kept(const double *C, long m1, long m2, const double *D, long r1, long r2,
                         double *out)
{
    const long d1 = m1 + 2 * r1, d2 = m2 + 2 * r2;
    const long k1 = 2 * r1 + 3, k2 = 2 * r2 + 3;
    const double mn = (double)(m1 * m2);
    double sT = 0.0;
    for (long i = 0; i < m1 * m2; ++i) sT += C[i];
    const double mT = sT / mn;
    double ssT = 0.0;
    for (long i = 0; i < m1 * m2; ++i) ssT += (C[i] - mT) * (C[i] - mT);
    const double denT = sqrt(ssT);
    double *den = (double *)malloc(sizeof(double) * (size_t)(k1 * k2));
    double dmax = 0.0;
    for (long b = 0; b < k2; ++b)
        for (long a = 0; a < k1; ++a) {
            const long s1 = a - 1, s2 = b - 1; /* template top-left at D(s1+1, s2+1) */
            double sTA = 0.0, sA = 0.0, sAA = 0.0;
            for (long j = 0; j < m2; ++j) {
                const long dj = s2 + j;
                if (dj < 0 || dj >= d2) continue;
                for (long i = 0; i < m1; ++i) {
                    const long di = s1 + i;
                    if (di < 0 || di >= d1) continue;
                    const double v = D[di + dj * d1];
                    sTA += C[i + j * m1] * v;
                    sA += v;
                    sAA += v * v;
                }
            }
            const double num = sTA - sA * sT / mn;
            double dA = sAA - sA * sA / mn;
            if (dA < 0) dA = 0;
            const double dn = denT * sqrt(dA);
            out[a + b * k1] = num;
            den[a + b * k1] = dn;
            if (dn > dmax) dmax = dn;
        }
    const double tol = sqrt(nextafter(dmax, INFINITY) - dmax);
    for (long i = 0; i < k1 * k2; ++i) out[i] = den[i] > tol ? out[i] / den[i] : 0.0;
    free(den);
}

/* normxcorr2(C, D), the whole map ((m1+d1-1) x (m2+d2-1), column-major):     */
/* every placement of the template over the zero-padded image, by direct    */
/* FP64 sums over the overlap (the public definition; no FFT, no running     */
/* sums).  Map entry (k, l), 0-based, puts the template's top-left at D(k -  */
/* m1 + 1, l - m2 + 1).  Used by funIntegerSearchMg.m:19,170.                */
void orc_normxcorr2_full(const double *C, long m1, long m2, const double *D, long d1, long d2,
                         double *out)
{
    const long k1 = m1 + d1 - 1, k2 = m2 + d2 - 1;
    const double mn = (double)(m1 * m2);
    double sT = 0.0;
    for (long i = 0; i < m1 * m2; ++i) sT += C[i];
    const double mT = sT / mn;
    double ssT = 0.0;
    for (long i = 0; i < m1 * m2; ++i) ssT += (C[i] - mT) * (C[i] - mT);
    const double denT = sqrt(ssT);
    double *den = (double *)malloc(sizeof(double) * (size_t)(k1 * k2));
    double dmax = 0.0;
    for (long l = 0; l < k2; ++l)
        for (long k = 0; k < k1; ++k) {
            const long s1 = k - m1 + 1, s2 = l - m2 + 1;
            double sTA = 0.0, sA = 0.0, sAA = 0.0;
            for (long j = 0; j < m2; ++j) {
                const long dj = s2 + j;
                if (dj < 0 || dj >= d2) continue;
                for (long i = 0; i < m1; ++i) {
                    const long di = s1 + i;
                    if (di < 0 || di >= d1) continue;
                    const double v = D[di + dj * d1];
                    sTA += C[i + j * m1] * v;
                    sA += v;
                    sAA += v * v;
                }
            }
            const double num = sTA - sA * sT / mn;
            double dA = sAA - sA * sA / mn;
            if (dA < 0) dA = 0;
            const double dn = denT * sqrt(dA);
            out[k + l * k1] = num;
            den[k + l * k1] = dn;
            if (dn > dmax) dmax = dn;
        }
    const double tol = sqrt(nextafter(dmax, INFINITY) - dmax);
    for (long i = 0; i < k1 * k2; ++i) out[i] = den[i] > tol ? out[i] / den[i] : 0.0;
    free(den);
}

/* ------------------------------------------------------------------------- */
/* funIntegerSearch.m:14-27 (ROI shrink) and :82-203 (whole-field loop).      */
/* ------------------------------------------------------------------------- */

int orc_integer_search(const double *f, const double *g, long H, long W, long r1, long r2,
                       const long gridx[2], const long gridy[2], long w, long step,
                       long max_pts, long *nx_out, long *ny_out, double *x0, double *y0,
                       double *u, double *v, double *ccmax)
{
    long gx0 = gridx[0], gx1 = gridx[1], gy0 = gridy[0], gy1 = gridy[1];
    while (gx0 - r1 - 0.5 * w < 1) ++gx0;
    while (gy0 - r2 - 0.5 * w < 1) ++gy0;
    while (gx1 + 1.5 * w + r1 > H) --gx1;
    while (gy1 + 1.5 * w + r2 > W) --gy1;
    /* XList = gx0 : step : gx1 - step (MATLAB colon, empty if end < start). */
    const long nx = (gx1 - step >= gx0) ? (gx1 - step - gx0) / step + 1 : 0;
    const long ny = (gy1 - step >= gy0) ? (gy1 - step - gy0) / step + 1 : 0;
    *nx_out = nx;
    *ny_out = ny;
    if (nx * ny > max_pts) return -1;
    const long m = w + 1, k1 = 2 * r1 + 3, k2 = 2 * r2 + 3;
    const long d1 = m + 2 * r1, d2 = m + 2 * r2;
#pragma omp parallel for schedule(dynamic)
    for (long t = 0; t < nx * ny; ++t) {
        const long ix = t % nx, iy = t / nx;
        const long ii = gx0 + ix * step, jj = gy0 + iy * step;
        double *C = (double *)malloc(sizeof(double) * (size_t)(m * m));
        double *D = (double *)malloc(sizeof(double) * (size_t)(d1 * d2));
        double *X = (double *)malloc(sizeof(double) * (size_t)(k1 * k2));
        for (long j = 0; j < m; ++j)
            for (long i = 0; i < m; ++i) C[i + j * m] = IDX(f, H, ii + i, jj + j);
        for (long j = 0; j < d2; ++j)
            for (long i = 0; i < d1; ++i) D[i + j * d1] = IDX(g, H, ii - r1 + i, jj - r2 + j);
        /* normxcorr2 rejects a constant template (error -> catch -> NaN). */
        int constant = 1;
        for (long i = 1; i < m * m; ++i)
            if (C[i] != C[0]) {
                constant = 0;
                break;
            }
        const long o = iy + ix * ny; /* outputs are ny x nx, rows = y */
        if (constant) {
            u[o] = v[o] = ccmax[o] = NAN;
        } else {
            double xp, yp, mf;
            orc_normxcorr2_kept(C, m, m, D, r1, r2, X);
            orc_findpeak(X, k1, k2, &xp, &yp, &mf);
            u[o] = yp - (double)(r1 + 2); /* zero_disp = ceil((2r+3)/2) */
            v[o] = xp - (double)(r2 + 2);
            ccmax[o] = mf;
        }
        x0[o] = (ii + ii + w) / 2.0;
        y0[o] = (jj + jj + w) / 2.0;
        free(C);
        free(D);
        free(X);
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Small dense solves.  MATLAB's mldivide takes a Cholesky factorisation for a */
/* symmetric matrix with a positive diagonal and falls back to LU with partial */
/* pivoting; the same order is used here.                                     */
/* ------------------------------------------------------------------------- */

static void solve_spd_or_lu(int n, const double *A_in, const double *b, double *x)
{
    double A[36], L[36];
    memcpy(A, A_in, sizeof(double) * (size_t)(n * n));
    int spd = 1;
    for (int i = 0; i < n && spd; ++i)
        if (!(A[i + i * n] > 0)) spd = 0;
    if (spd) {
        memset(L, 0, sizeof(L));
        for (int j = 0; j < n && spd; ++j) {
            double d = A[j + j * n];
            for (int k = 0; k < j; ++k) d -= L[j + k * n] * L[j + k * n];
            if (!(d > 0)) {
                spd = 0;
                break;
            }
            const double ljj = sqrt(d);
            L[j + j * n] = ljj;
            for (int i = j + 1; i < n; ++i) {
                double s = A[i + j * n];
                for (int k = 0; k < j; ++k) s -= L[i + k * n] * L[j + k * n];
                L[i + j * n] = s / ljj;
            }
        }
        if (spd) {
            double y[6];
            for (int i = 0; i < n; ++i) {
                double s = b[i];
                for (int k = 0; k < i; ++k) s -= L[i + k * n] * y[k];
                y[i] = s / L[i + i * n];
            }
            for (int i = n - 1; i >= 0; --i) {
                double s = y[i];
                for (int k = i + 1; k < n; ++k) s -= L[k + i * n] * x[k];
                x[i] = s / L[i + i * n];
            }
            return;
        }
    }
    /* LU with partial pivoting. */
    double bb[6];
    memcpy(bb, b, sizeof(double) * (size_t)n);
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = fabs(A[k + k * n]);
        for (int i = k + 1; i < n; ++i)
            if (fabs(A[i + k * n]) > best) {
                best = fabs(A[i + k * n]);
                p = i;
            }
        if (p != k) {
            for (int j = 0; j < n; ++j) {
                const double t = A[k + j * n];
                A[k + j * n] = A[p + j * n];
                A[p + j * n] = t;
            }
            const double t = bb[k];
            bb[k] = bb[p];
            bb[p] = t;
        }
        for (int i = k + 1; i < n; ++i) {
            const double l = A[i + k * n] / A[k + k * n];
            A[i + k * n] = l;
            for (int j = k + 1; j < n; ++j) A[i + j * n] -= l * A[k + j * n];
            bb[i] -= l * bb[k];
        }
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = bb[i];
        for (int j = i + 1; j < n; ++j) s -= A[i + j * n] * x[j];
        x[i] = s / A[i + i * n];
    }
}

/* Inverse-compositional update W(P) <- W(P) o W(dP)^-1 with the closed-form
 * inverse of funICGNQuadtree.m:393-411.  Returns 0 if det(dP) == 0. */
static int ic_compose(double P[6], const double d[6])
{
    const double det = (1 + d[0]) * (1 + d[3]) - d[1] * d[2];
    if (!(det != 0)) return 0;
    const double q1 = (-d[0] - d[0] * d[3] + d[1] * d[2]) / det;
    const double q2 = -d[1] / det;
    const double q3 = -d[2] / det;
    const double q4 = (-d[3] - d[0] * d[3] + d[1] * d[2]) / det;
    const double q5 = (-d[4] - d[3] * d[4] + d[2] * d[5]) / det;
    const double q6 = (-d[5] - d[0] * d[5] + d[1] * d[4]) / det;
    /* M = [1+P1 P3 P5; P2 1+P4 P6; 0 0 1] * [1+q1 q3 q5; q2 1+q4 q6; 0 0 1] */
    const double a11 = 1 + P[0], a12 = P[2], a13 = P[4];
    const double a21 = P[1], a22 = 1 + P[3], a23 = P[5];
    const double b11 = 1 + q1, b12 = q3, b13 = q5;
    const double b21 = q2, b22 = 1 + q4, b23 = q6;
    const double m11 = a11 * b11 + a12 * b21;
    const double m12 = a11 * b12 + a12 * b22;
    const double m13 = a11 * b13 + a12 * b23 + a13;
    const double m21 = a21 * b11 + a22 * b21;
    const double m22 = a21 * b12 + a22 * b22;
    const double m23 = a21 * b13 + a22 * b23 + a23;
    P[0] = m11 - 1;
    P[1] = m21;
    P[2] = m12;
    P[3] = m22 - 1;
    P[4] = m13;
    P[5] = m23;
    return 1;
}

/* mean and sqrt((n-1)*var) over the entries with |a| > 1e-10 (masked) or all
 * entries (whole), funICGNQuadtree.m:119-120 / funICGN.m:93-94. */
static void subset_stats(const double *a, long n, int whole, double *mean, double *bottom)
{
    long cnt = 0;
    double s = 0.0;
    for (long i = 0; i < n; ++i)
        if (whole || fabs(a[i]) > 1e-10) {
            s += a[i];
            ++cnt;
        }
    if (cnt == 0) {
        *mean = NAN;
        *bottom = NAN;
        return;
    }
    const double m = s / (double)cnt;
    double ss = 0.0;
    for (long i = 0; i < n; ++i)
        if (whole || fabs(a[i]) > 1e-10) ss += (a[i] - m) * (a[i] - m);
    *mean = m;
    *bottom = cnt == 1 ? 0.0 : sqrt(ss);
}

/* A window [xa..xb] x [ya..yb] (1-based, inclusive) can be read from the image,
 * the mask and the gradient arrays (Df index = pixel - 3 in 1..H-6). */
static int window_ok(const orc_frame *fr, double xa, double xb, double ya, double yb)
{
    if (xa != floor(xa) || ya != floor(ya) || xb != floor(xb) || yb != floor(yb)) return 0;
    if (!(xa >= 4 && ya >= 4 && xb <= fr->H - 3 && yb <= fr->W - 3)) return 0;
    return 1;
}

typedef struct {
    long nx, ny;   /* window extent (wx+1, wy+1) */
    long xa, ya;   /* first pixel (1-based) */
    double *f, *m, *fx, *fy;
} window_t;

static void window_load(const orc_frame *fr, window_t *w, int mask_f)
{
    const long H = fr->H, Hd = fr->H - 6;
    for (long j = 0; j < w->ny; ++j)
        for (long i = 0; i < w->nx; ++i) {
            const long x = w->xa + i, y = w->ya + j;
            const long k = i + j * w->nx;
            const double mk = IDX(fr->mask, H, x, y);
            w->m[k] = mk;
            w->f[k] = IDX(fr->f, H, x, y) * (mask_f ? mk : 1.0);
            w->fx[k] = fr->dfdx[(x - 4) + (y - 4) * Hd];
            w->fy[k] = fr->dfdy[(x - 4) + (y - 4) * Hd];
        }
}

static void window_alloc(window_t *w, long nx, long ny)
{
    w->nx = nx;
    w->ny = ny;
    w->f = (double *)malloc(sizeof(double) * (size_t)(nx * ny));
    w->m = (double *)malloc(sizeof(double) * (size_t)(nx * ny));
    w->fx = (double *)malloc(sizeof(double) * (size_t)(nx * ny));
    w->fy = (double *)malloc(sizeof(double) * (size_t)(nx * ny));
}

static void window_free(window_t *w)
{
    free(w->f);
    free(w->m);
    free(w->fx);
    free(w->fy);
}

/* Warped sample positions u22, v22 (funICGNQuadtree.m:149-152) for window
 * offsets dx = i - wx/2, dy = j - wy/2.  Returns 1 if any lies outside
 * [3, H-2] x [3, W-2] (:154-157). */
static int warp_coords(const orc_frame *fr, const double P[6], double x0, double y0, long nx,
                       long ny, double *u22, double *v22)
{
    int out = 0;
    const double hx = (nx - 1) / 2.0, hy = (ny - 1) / 2.0;
    for (long j = 0; j < ny; ++j)
        for (long i = 0; i < nx; ++i) {
            const double dx = i - hx, dy = j - hy;
            const long k = i + j * nx;
            u22[k] = (1 + P[0]) * dx + P[2] * dy + (x0 + P[4]);
            v22[k] = P[1] * dx + (1 + P[3]) * dy + (y0 + P[5]);
            if (u22[k] < 3 || u22[k] > fr->H - 2 || v22[k] < 3 || v22[k] > fr->W - 2) out = 1;
        }
    return out;
}

/* funICGNQuadtree.m:1-451 (plain = 0) / funICGN.m:1-323 (plain = 1). */
int orc_icgn6(const orc_frame *fr, const double U0[2], double x0, double y0, long w,
              double tol, int plain, double P[6], double H21[21])
{
    const long w0 = w;
    const int max_it = 100;
    double Hm[36];
    memset(Hm, 0, sizeof(Hm));
    for (int k = 0; k < 21; ++k) H21[k] = 0.0;
    P[0] = P[1] = P[2] = P[3] = 0.0;
    P[4] = U0[0];
    P[5] = U0[1];

    if (!window_ok(fr, x0 - w / 2.0, x0 + w / 2.0, y0 - w / 2.0, y0 + w / 2.0)) goto error;

    window_t win;
    window_alloc(&win, w + 1, w + 1);
    win.xa = (long)(x0 - w / 2);
    win.ya = (long)(y0 - w / 2);
    window_load(fr, &win, !plain);

    long n = win.nx * win.ny;
    /* masked-pixel count: zeros of tempf (:61) / zeros of DfDx (funICGN.m:54) */
    long count = 0;
    for (long k = 0; k < n; ++k) count += plain ? (win.fx[k] == 0) : (win.f[k] == 0);
    const double cap = plain ? 0.5 * (w + 1) * (w + 1) : 0.4 * (w + 1) * (w + 1);
    if (!(count < cap)) {
        window_free(&win);
        P[0] = P[1] = P[2] = P[3] = 0.0;
        P[4] = U0[0];
        P[5] = U0[1];
        return ORC_ST_MASKED; /* H = 0 */
    }
    int grow = plain ? (count > 0.1 * (w + 1) * (w + 1)) : (count > 0);
    if (grow) {
        if (plain)
            w = lround(sqrt(2.0) * w); /* MATLAB round of sqrt(2)*w */
        else
            w = 2 * (long)ceil(sqrt((double)((w + 1) * (w + 1)) / ((w + 1) * (w + 1) - count)) * w / 2.0);
        window_free(&win);
        if (!window_ok(fr, x0 - w / 2.0, x0 + w / 2.0, y0 - w / 2.0, y0 + w / 2.0)) goto error;
        window_alloc(&win, w + 1, w + 1);
        win.xa = (long)(x0 - w / 2);
        win.ya = (long)(y0 - w / 2);
        window_load(fr, &win, !plain);
        n = win.nx * win.ny;
    }
    unsigned char *bw = (unsigned char *)malloc((size_t)n);
    unsigned char *sel = (unsigned char *)malloc((size_t)n);
    const long seed = (w + 1) / 2; /* floor((winsize+1)/2) */
    if (!plain) {
        /* :83-88 — 4-connected region of the window mask through the seed. */
        for (long k = 0; k < n; ++k) bw[k] = win.m[k] != 0;
        orc_bwselect(bw, win.nx, win.ny, seed, seed, 4, sel);
        for (long k = 0; k < n; ++k)
            if (!sel[k]) win.f[k] = win.fx[k] = win.fy[k] = 0.0;
    }
    /* :91-103 — Hessian of the affine steepest-descent images. */
    {
        const double hx = w / 2.0;
        double Hs[21] = {0};
        for (long j = 0; j < win.ny; ++j)
            for (long i = 0; i < win.nx; ++i) {
                const long k = i + j * win.nx;
                const double dx = i - hx, dy = j - hx;
                const double J[6] = {win.fx[k] * dx, win.fy[k] * dx, win.fx[k] * dy,
                                     win.fy[k] * dy, win.fx[k], win.fy[k]};
                int t = 0;
                for (int c = 0; c < 6; ++c)
                    for (int r = c; r < 6; ++r) Hs[t++] += J[r] * J[c];
            }
        int t = 0;
        for (int c = 0; c < 6; ++c)
            for (int r = c; r < 6; ++r) {
                Hm[r + c * 6] = Hm[c + r * 6] = Hs[t];
                H21[t] = Hs[t];
                ++t;
            }
    }
    double meanf, bottomf;
    subset_stats(win.f, n, plain, &meanf, &bottomf);

    double *u22 = (double *)malloc(sizeof(double) * (size_t)n);
    double *v22 = (double *)malloc(sizeof(double) * (size_t)n);
    double *gv = (double *)malloc(sizeof(double) * (size_t)n);
    double normNew = 1, normAbs = 1, normInit = 0;
    long sum_bw2 = 0;
    int it = 0, status;
    while (it <= max_it && normNew > tol && normAbs > tol) {
        ++it;
        if (!plain && it > 1 && count > 0) {
            /* :139-145 — the window is re-sized from the previous g region but
             * tempf/DfDx are not reloaded; a changed size makes the residual
             * at :363 throw (caught as status -1), unless the new grid leaves
             * the image first (:154-157, status 101). */
            const double wn = 2.0 * ceil(sqrt((double)((w0 + 1) * (w0 + 1)) / (double)sum_bw2) * w / 2.0);
            if (wn != (double)w) {
                if (!isfinite(wn) || x0 - wn / 2 < 1 || x0 + wn / 2 > fr->H || y0 - wn / 2 < 1 ||
                    y0 + wn / 2 > fr->W)
                    goto error_loop;
                const long wl = (long)wn;
                double *uu = (double *)malloc(sizeof(double) * (size_t)((wl + 1) * (wl + 1)));
                double *vv = (double *)malloc(sizeof(double) * (size_t)((wl + 1) * (wl + 1)));
                const int out = warp_coords(fr, P, x0, y0, wl + 1, wl + 1, uu, vv);
                free(uu);
                free(vv);
                if (out) {
                    normNew = 1e6;
                    break;
                }
                goto error_loop;
            }
        }
        if (warp_coords(fr, P, x0, y0, win.nx, win.ny, u22, v22)) {
            normNew = 1e6;
            break;
        }
        long zc = 0;
        int has_nan = 0;
        for (long k = 0; k < n; ++k) {
            gv[k] = bicubic_sample(fr->g, fr->H, fr->W, v22[k], u22[k], NULL, NULL);
            if (isnan(gv[k])) has_nan = 1;
            zc += gv[k] == 0;
        }
        if (!plain) {
            if (has_nan) goto error_loop; /* logical(NaN) throws (:163) */
            if (count > 0 || zc > 0) {
                /* :171-178 — 8-connected region of nonzero g through the seed,
                 * or of the window mask if that region is the whole window. */
                for (long k = 0; k < n; ++k) bw[k] = gv[k] != 0;
                orc_bwselect(bw, win.nx, win.ny, seed, seed, 8, sel);
                int all = 1;
                for (long k = 0; k < n; ++k)
                    if (!sel[k]) {
                        all = 0;
                        break;
                    }
                if (all) {
                    for (long k = 0; k < n; ++k) bw[k] = win.m[k] != 0;
                    orc_bwselect(bw, win.nx, win.ny, seed, seed, 8, sel);
                }
                sum_bw2 = 0;
                for (long k = 0; k < n; ++k) {
                    if (!sel[k]) gv[k] = 0.0;
                    sum_bw2 += sel[k];
                }
            }
        }
        double meang, bottomg;
        subset_stats(gv, n, plain, &meang, &bottomg);
        /* :362-371 */
        double b[6] = {0};
        const double hx = w / 2.0;
        for (long j = 0; j < win.ny; ++j)
            for (long i = 0; i < win.nx; ++i) {
                const long k = i + j * win.nx;
                const double r = (win.f[k] - meanf) / bottomf - (gv[k] - meang) / bottomg;
                const double dx = i - hx, dy = j - hx;
                b[0] += dx * win.fx[k] * r;
                b[1] += dx * win.fy[k] * r;
                b[2] += dy * win.fx[k] * r;
                b[3] += dy * win.fy[k] * r;
                b[4] += win.fx[k] * r;
                b[5] += win.fy[k] * r;
            }
        double nb = 0.0;
        for (int k = 0; k < 6; ++k) {
            b[k] *= bottomf;
            nb += b[k] * b[k];
        }
        normNew = sqrt(nb);
        normAbs = normNew;
        if (it == 1) normInit = normNew;
        normNew = normInit > tol ? normNew / normInit : 0.0;
        if (normNew < tol || normAbs < tol) break;
        double d[6], nbv[6];
        for (int k = 0; k < 6; ++k) nbv[k] = -b[k];
        solve_spd_or_lu(6, Hm, nbv, d);
        if (!ic_compose(P, d)) break;
    }
    if (!(normNew < tol || normAbs < tol)) it = max_it + 1;
    if (isnan(normNew)) it = max_it + 1;
    if (!plain) {
        double sa = 0.0;
        for (long k = 0; k < n; ++k) sa += fabs(win.f[k]);
        if (sa < 1e-6) it = max_it + 3;
    }
    status = it;
    free(u22);
    free(v22);
    free(gv);
    free(bw);
    free(sel);
    window_free(&win);
    return status;

error_loop:
    free(u22);
    free(v22);
    free(gv);
    free(bw);
    free(sel);
    window_free(&win);
error:
    for (int k = 0; k < 6; ++k) P[k] = NAN;
    for (int k = 0; k < 21; ++k) H21[k] = 0.0;
    return ORC_ST_ERROR;
}

/* funICGN_Subpb1.m:1-474 (Gauss-Newton branch). */
int orc_icgn2_al(const orc_frame *fr, double x0, double y0, long wx, long wy, double mu,
                 const double vdual[2], const double UOld[2], const double FOld[4], double tol,
                 double U[2])
{
    const int max_it = 100;
    double P[6] = {FOld[0], FOld[1], FOld[2], FOld[3], UOld[0], UOld[1]};
    U[0] = UOld[0];
    U[1] = UOld[1];
    if (!window_ok(fr, x0 - wx / 2.0, x0 + wx / 2.0, y0 - wy / 2.0, y0 + wy / 2.0)) goto error;
    window_t win;
    window_alloc(&win, wx + 1, wy + 1);
    win.xa = (long)(x0 - wx / 2);
    win.ya = (long)(y0 - wy / 2);
    window_load(fr, &win, 1);
    long n = win.nx * win.ny;
    long count = 0;
    for (long k = 0; k < n; ++k) count += win.f[k] == 0;
    if (!(count < 0.4 * (wx + 1) * (wy + 1))) {
        window_free(&win);
        return ORC_ST_MASKED;
    }
    if (count > 0) {
        /* :78-79 — note wy uses the already-updated wx. */
        wx = 2 * (long)ceil(sqrt((double)((wx + 1) * (wy + 1)) / ((wx + 1) * (wy + 1) - count)) * wx / 2.0);
        wy = 2 * (long)ceil(sqrt((double)((wx + 1) * (wy + 1)) / ((wx + 1) * (wy + 1) - count)) * wy / 2.0);
        window_free(&win);
        if (!window_ok(fr, x0 - wx / 2.0, x0 + wx / 2.0, y0 - wy / 2.0, y0 + wy / 2.0)) goto error;
        window_alloc(&win, wx + 1, wy + 1);
        win.xa = (long)(x0 - wx / 2);
        win.ya = (long)(y0 - wy / 2);
        window_load(fr, &win, 1);
        n = win.nx * win.ny;
    }
    /* bwselect(BW, c, r): c = floor((wx+1)/2) is a column (dim-2) index and
     * r = floor((wy+1)/2) a row (dim-1) index (:93, :160, :165). */
    const long sc = (wx + 1) / 2, sr = (wy + 1) / 2;
    unsigned char *bw = (unsigned char *)malloc((size_t)n);
    unsigned char *sel = (unsigned char *)malloc((size_t)n);
    for (long k = 0; k < n; ++k) bw[k] = win.m[k] != 0;
    orc_bwselect(bw, win.nx, win.ny, sc, sr, 4, sel);
    for (long k = 0; k < n; ++k)
        if (!sel[k]) win.f[k] = win.fx[k] = win.fy[k] = 0.0;
    double h11 = 0, h12 = 0, h22 = 0;
    for (long k = 0; k < n; ++k) {
        h11 += win.fx[k] * win.fx[k];
        h12 += win.fx[k] * win.fy[k];
        h22 += win.fy[k] * win.fy[k];
    }
    double meanf, bottomf;
    subset_stats(win.f, n, 0, &meanf, &bottomf);
    double H2[4] = {h11 * 2 / (bottomf * bottomf) + mu, h12 * 2 / (bottomf * bottomf),
                    h12 * 2 / (bottomf * bottomf), h22 * 2 / (bottomf * bottomf) + mu};

    double *u22 = (double *)malloc(sizeof(double) * (size_t)n);
    double *v22 = (double *)malloc(sizeof(double) * (size_t)n);
    double *gv = (double *)malloc(sizeof(double) * (size_t)n);
    double normNew = 1, normAbs = 1, normInit = 0;
    int it = 0, status;
    while (it <= 100 && normNew > tol && normAbs > tol) {
        ++it;
        if (warp_coords(fr, P, x0, y0, win.nx, win.ny, u22, v22)) {
            normNew = 1e6;
            break;
        }
        long zc = 0;
        int has_nan = 0;
        for (long k = 0; k < n; ++k) {
            gv[k] = bicubic_sample(fr->g, fr->H, fr->W, v22[k], u22[k], NULL, NULL);
            if (isnan(gv[k])) has_nan = 1;
            zc += gv[k] == 0;
        }
        if (has_nan) goto error_loop; /* logical(NaN) throws (:153) */
        if (count > 0 || zc > 0) {
            /* :160-190 — region of nonzero g (4-connected), or the window mask
             * (8-connected) when that is the whole window; f, grad f, H and the
             * f statistics are rebuilt on the region every iteration. */
            for (long k = 0; k < n; ++k) bw[k] = gv[k] != 0;
            orc_bwselect(bw, win.nx, win.ny, sc, sr, 4, sel);
            int all = 1;
            for (long k = 0; k < n; ++k)
                if (!sel[k]) {
                    all = 0;
                    break;
                }
            if (all) {
                for (long k = 0; k < n; ++k) bw[k] = win.m[k] != 0;
                orc_bwselect(bw, win.nx, win.ny, sc, sr, 8, sel);
            }
            window_load(fr, &win, 1);
            for (long k = 0; k < n; ++k)
                if (!sel[k]) gv[k] = win.f[k] = win.fx[k] = win.fy[k] = 0.0;
            h11 = h12 = h22 = 0;
            for (long k = 0; k < n; ++k) {
                h11 += win.fx[k] * win.fx[k];
                h12 += win.fx[k] * win.fy[k];
                h22 += win.fy[k] * win.fy[k];
            }
            subset_stats(win.f, n, 0, &meanf, &bottomf);
            H2[0] = h11 * 2 / (bottomf * bottomf) + mu;
            H2[1] = H2[2] = h12 * 2 / (bottomf * bottomf);
            H2[3] = h22 * 2 / (bottomf * bottomf) + mu;
        }
        double meang, bottomg;
        subset_stats(gv, n, 0, &meang, &bottomg);
        double b1 = 0, b2 = 0;
        for (long k = 0; k < n; ++k) {
            const double r = (win.f[k] - meanf) / bottomf - (gv[k] - meang) / bottomg;
            b1 += win.fx[k] * r;
            b2 += win.fy[k] * r;
        }
        b1 *= bottomf;
        b2 *= bottomf;
        /* :379 */
        const double tb[2] = {b1 * 2 / (bottomf * bottomf) + mu * (P[4] - UOld[0] - vdual[0]),
                              b2 * 2 / (bottomf * bottomf) + mu * (P[5] - UOld[1] - vdual[1])};
        /* :391-394 */
        const double dl = 1e-3 * (H2[0] > H2[3] ? H2[0] : H2[3]);
        const double tH[4] = {H2[0] + dl, H2[1], H2[2], H2[3] + dl};
        const double ntb[2] = {-tb[0], -tb[1]};
        double d56[2];
        solve_spd_or_lu(2, tH, ntb, d56);
        normNew = sqrt(tb[0] * tb[0] + tb[1] * tb[1]);
        normAbs = normNew;
        if (it == 1) normInit = normNew;
        normNew = normInit > tol ? normNew / normInit : 0.0;
        if (normNew < tol || normAbs < tol) break;
        const double d[6] = {0, 0, 0, 0, d56[0], d56[1]};
        if (!ic_compose(P, d)) break;
    }
    if (!(normNew < tol || normAbs < tol)) it = max_it + 1;
    if (isnan(normNew)) it = max_it + 1;
    {
        double sa = 0.0;
        for (long k = 0; k < n; ++k) sa += fabs(win.f[k]);
        if (sa < 1e-6) it = max_it + 3;
    }
    status = it;
    U[0] = P[4];
    U[1] = P[5];
    free(u22);
    free(v22);
    free(gv);
    free(bw);
    free(sel);
    window_free(&win);
    return status;

error_loop:
    free(u22);
    free(v22);
    free(gv);
    free(bw);
    free(sel);
    window_free(&win);
error:
    U[0] = U[1] = NAN;
    return ORC_ST_ERROR;
}

/* ------------------------------------------------------------------------- */
/* Batch drivers (LocalICGNQuadtree.m:63-147, Subpb1Quadtree.m:70-169).        */
/* ------------------------------------------------------------------------- */

static double node_coord(double c, int plain) { return plain ? c : round(c); }

void orc_local_icgn_batch(const orc_frame *fr, const double *U0, const double *coords,
                          const long *winsize, long N, double tol, int plain, int nthreads,
                          double *U, double *F, double *Hpar, int *status)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic)
    for (long j = 0; j < N; ++j) {
        double P[6], H21[21];
        const double x0 = node_coord(coords[j], plain), y0 = node_coord(coords[j + N], plain);
        const int st = orc_icgn6(fr, &U0[2 * j], x0, y0, winsize[j], tol, plain, P, H21);
        status[j] = st;
        U[2 * j] = P[4];
        U[2 * j + 1] = P[5];
        for (int k = 0; k < 4; ++k) F[4 * j + k] = P[k];
        for (int k = 0; k < 21; ++k) Hpar[j + k * N] = H21[k];
    }
}

void orc_subpb1_batch(const orc_frame *fr, const double *UOld, const double *FOld,
                      const double *vdual, const double *coords, const long *wx,
                      const long *wy, long N, double mu, double tol, int plain, int nthreads,
                      double *U, int *status)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic)
    for (long j = 0; j < N; ++j) {
        const double x0 = node_coord(coords[j], plain), y0 = node_coord(coords[j + N], plain);
        status[j] = orc_icgn2_al(fr, x0, y0, wx[j], wy[j], mu, &vdual[2 * j], &UOld[2 * j],
                                 &FOld[4 * j], tol, &U[2 * j]);
    }
}

/* ------------------------------------------------------------------------- */
/* ZNSSD mismatch and its gradient (the adjoint of the warp-gather).           */
/* dC/dg_p = -(2/bg) (r_p - rbar - ghat_p * sum_q r_q ghat_q),  r = fhat-ghat;  */
/* dg/dP via the derivative of the Keys kernel and du/dP = [dx 0 dy 0 1 0],    */
/* dv/dP = [0 dx 0 dy 0 1].                                                    */
/* ------------------------------------------------------------------------- */

double orc_znssd(const orc_frame *fr, const double *P, const double *coords, const long *winsize,
                 long N, double *cost, double *grad, int *status_out)
{
    double total = 0.0;
#pragma omp parallel for schedule(dynamic) reduction(+ : total)
    for (long j = 0; j < N; ++j) {
        const long w = winsize[j];
        const long nx = w + 1, n = nx * nx;
        const double x0 = round(coords[j]), y0 = round(coords[j + N]);
        const double *p = &P[6 * j];
        double c = 0.0, gr[6] = {0};
        int st = 0;
        if (!window_ok(fr, x0 - w / 2.0, x0 + w / 2.0, y0 - w / 2.0, y0 + w / 2.0)) st = 1;
        double *fv = NULL, *gv = NULL, *gu = NULL, *gvv = NULL, *u22 = NULL, *v22 = NULL;
        if (!st) {
            fv = (double *)malloc(sizeof(double) * (size_t)n);
            gv = (double *)malloc(sizeof(double) * (size_t)n);
            gu = (double *)malloc(sizeof(double) * (size_t)n);
            gvv = (double *)malloc(sizeof(double) * (size_t)n);
            u22 = (double *)malloc(sizeof(double) * (size_t)n);
            v22 = (double *)malloc(sizeof(double) * (size_t)n);
            if (warp_coords(fr, p, x0, y0, nx, nx, u22, v22)) st = 1;
        }
        if (!st) {
            for (long jj = 0; jj < nx; ++jj)
                for (long ii = 0; ii < nx; ++ii) {
                    const long k = ii + jj * nx;
                    fv[k] = IDX(fr->f, fr->H, (long)x0 - w / 2 + ii, (long)y0 - w / 2 + jj);
                    double dX, dY;
                    gv[k] = bicubic_sample(fr->g, fr->H, fr->W, v22[k], u22[k], &dX, &dY);
                    gu[k] = dY; /* along dim 1 (u22) */
                    gvv[k] = dX; /* along dim 2 (v22) */
                }
            double mf, bf, mg, bg;
            subset_stats(fv, n, 1, &mf, &bf);
            subset_stats(gv, n, 1, &mg, &bg);
            double rbar = 0.0, srg = 0.0;
            for (long k = 0; k < n; ++k) {
                const double fh = (fv[k] - mf) / bf, gh = (gv[k] - mg) / bg;
                const double r = fh - gh;
                c += r * r;
                rbar += r;
                srg += r * gh;
            }
            rbar /= (double)n;
            if (grad) {
                const double hx = w / 2.0;
                for (long jj = 0; jj < nx; ++jj)
                    for (long ii = 0; ii < nx; ++ii) {
                        const long k = ii + jj * nx;
                        const double fh = (fv[k] - mf) / bf, gh = (gv[k] - mg) / bg;
                        const double dCdg = -(2.0 / bg) * ((fh - gh) - rbar - gh * srg);
                        const double dx = ii - hx, dy = jj - hx;
                        const double au = dCdg * gu[k], av = dCdg * gvv[k];
                        gr[0] += au * dx;
                        gr[1] += av * dx;
                        gr[2] += au * dy;
                        gr[3] += av * dy;
                        gr[4] += au;
                        gr[5] += av;
                    }
            }
        }
        free(fv);
        free(gv);
        free(gu);
        free(gvv);
        free(u22);
        free(v22);
        if (cost) cost[j] = st ? 0.0 : c;
        if (grad)
            for (int k = 0; k < 6; ++k) grad[6 * j + k] = st ? 0.0 : gr[k];
        if (status_out) status_out[j] = st;
        total += st ? 0.0 : c;
    }
    return total;
}
