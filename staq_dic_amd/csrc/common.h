// Device-side types and helpers shared by the STAQ-DIC HIP kernels (gfx950).
//
// Layout conventions (identical to the reference after its transpose on load,
// func_quadtree/ReadImageQuadtree.m:78): images are column-major with dim 1 = x
// (H entries, contiguous) and dim 2 = y (W entries); coordinates are 1-based.
// Gradient arrays are (H-6) x (W-6), element (i, j) belongs to pixel (i+3, j+3)
// (func/funImgGradient.m:53-78).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace staq {

// One reference/deformed frame pair resident in HBM (FP32).
struct FrameDev {
    const float *f;     // reference image (normalised), H x W
    const float *g;     // deformed image, H x W
    const float *mask;  // reference mask (0/1), H x W
    const float *dfdx;  // (H-6) x (W-6), masked
    const float *dfdy;  // (H-6) x (W-6), masked
    int H, W;
};

// Status codes of the reference (funICGNQuadtree.m:422-442, LocalICGNQuadtree.m:92-97),
// and two internal codes that never leave a local step: ST_DEFER (the window
// or the warped support outgrew the kernel) and ST_EXACT (a stopping test
// fell inside the guard band around tol; the subset is solved again with
// FP64 arithmetic throughout, so that its iteration count is the one an FP64
// evaluation gives).
enum : int {
    ST_EXACT = -3,
    ST_DEFER = -2,
    ST_NOTCONV = 101,
    ST_MASKED = 102,
    ST_FEATURELESS = 103,
    ST_ERROR = -1
};

// Stopping test inside the guard band: x (a squared norm) within a relative
// band of width g2 (> 0) around tol2 on either side.  FP32 sums place |b|
// within a small relative error of its FP64 value, so only a test this close
// to tol can come out differently.
__device__ __forceinline__ bool near_tol(double x, double tol2, double g2)
{
    return g2 > 0 && x > tol2 * (1.0 - g2) && x < tol2 * (1.0 + g2);
}

// Spread max L_cc^2 / min L_cc^2 of the Cholesky pivots of the 6x6 Hessian,
// a lower bound of its condition number.  Past kPivotSpread the FP32 sums of
// H and b can move the solution of an ill-conditioned window (small windows
// with weak texture along one direction) by more than 1e-3 px from an FP64
// evaluation, so such subsets take the FP64 re-solve.
constexpr double kPivotSpread = 3e3;
__device__ __forceinline__ bool pivots_spread(const double d[6])
{
    double lo = d[0] * d[0], hi = lo;
#pragma unroll
    for (int q = 1; q < 6; ++q) {
        lo = fmin(lo, d[q] * d[q]);
        hi = fmax(hi, d[q] * d[q]);
    }
    return !(hi <= kPivotSpread * lo);
}

// A warped corner within 1e-3 px of either bound of the image test
// (funICGNQuadtree.m:154-157), where the FP32 sums' effect on P could decide it.
__device__ __forceinline__ bool near_edge(double u, double lo, double hi)
{
    return fabs(u - lo) < 1e-3 || fabs(u - hi) < 1e-3;
}

// Warp coordinates for the Keys gather, split so that the fraction that
// decides the weights keeps FP32 precision.  Window pixel (x0 + dx, y0 + dy)
// warps to u = (x0 + dx + floor(P5)) + s_u with
//   s_u = P1 dx + P3 dy + (P5 - floor(P5)),
// where x0 + dx is an integer (the window's corners are integers) and s_u is
// a few pixels at most; s_u is formed in FP32, so its rounding error is that
// of a number of magnitude |P1 dx + P3 dy| + 1 rather than of the tile-local
// or absolute coordinate.  floor(u) = (x0 + dx + floor(P5)) + floor(s_u)
// exactly, and the Keys fraction is s_u - floor(s_u), taken with the
// hardware fract (v_fract_f32: the same value, clamped below 1).  v likewise.
struct SplitWarp {
    float p1, p2, p3, p4;  // P[0..3] in FP32
    float fu, fv;          // fractional parts of P5, P6 (in [0, 1])
    // cu = x0 - tu0 + floor(P5) - 1: the 0-based tile index of the first
    // tap of window offset dx is dx + cu + floor(s_u) (cu is integer- or
    // half-integer-valued like -dx, so the sum is exact in FP32); cv likewise
    float cu, cv;
};

__device__ __forceinline__ SplitWarp split_warp(const double P[6], double x0, double y0, int tu0,
                                                int tv0)
{
    SplitWarp s;
    s.p1 = (float)P[0];
    s.p2 = (float)P[1];
    s.p3 = (float)P[2];
    s.p4 = (float)P[3];
    const double ip5 = floor(P[4]), ip6 = floor(P[5]);
    s.fu = (float)(P[4] - ip5);
    s.fv = (float)(P[5] - ip6);
    s.cu = (float)(x0 - tu0 + ip5 - 1.0);
    s.cv = (float)(y0 - tv0 + ip6 - 1.0);
    return s;
}


constexpr int kWave = 64;

// Loads through the global address space.  Pointers read from a struct in
// memory are generic (flat) to the compiler, and flat loads share the LDS
// wait counter, so every LDS wait would also wait for HBM.
template <typename T>
__device__ __forceinline__ T gld(const T *p, size_t i)
{
    return ((const __attribute__((address_space(1))) T *)p)[i];
}

// One frame descriptor, field by field through the global address space (the
// loads are then known to be the same in every lane when slot is).
__device__ __forceinline__ FrameDev load_frame(const FrameDev *frames, int slot)
{
    const __attribute__((address_space(1))) FrameDev *q =
        (const __attribute__((address_space(1))) FrameDev *)frames + slot;
    FrameDev fr;
    fr.f = q->f;
    fr.g = q->g;
    fr.mask = q->mask;
    fr.dfdx = q->dfdx;
    fr.dfdy = q->dfdy;
    fr.H = q->H;
    fr.W = q->W;
    return fr;
}

__device__ __forceinline__ float fidx(const float *a, int H, int x, int y)
{
    // 1-based (x, y) in an H-row column-major array
    return gld(a, (size_t)((x - 1) + (y - 1) * H));  // images stay below 2^31 pixels
}

// One DPP move of a double (both 32-bit halves); lanes without a source get 0.
template <int ctrl>
__device__ __forceinline__ double dpp_f64(double v)
{
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)bits, ctrl, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(bits >> 32), ctrl, 0xf, 0xf, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <int ctrl>
__device__ __forceinline__ unsigned long long dpp_u64(unsigned long long v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, (int)v, ctrl, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(v >> 32), ctrl, 0xf, 0xf, true);
    return ((unsigned long long)(unsigned int)hi << 32) | (unsigned int)lo;
}

__device__ __forceinline__ double readlane_f64(double v, int l)
{
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)bits, l);
    const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// FP64 wavefront sum with the same DPP scheme, result in every lane.
__device__ __forceinline__ double wave_sum_dpp_f64(double v)
{
    v += dpp_f64<0x111>(v);
    v += dpp_f64<0x112>(v);
    v += dpp_f64<0x114>(v);
    v += dpp_f64<0x118>(v);
    return (readlane_f64(v, 15) + readlane_f64(v, 31)) + (readlane_f64(v, 47) + readlane_f64(v, 63));
}

// Sum over the 64 lanes of a wavefront, result in every lane: DPP row shifts
// build each 16-lane row's sum in its last lane, two row broadcasts add the
// rows into lane 63 ((r3 + r2) + (r1 + r0)), one readlane returns it.
__device__ __forceinline__ float wave_sum_dpp(float v)
{
    int x = __float_as_int(v);
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, true));  // row_shr:1
    x = __float_as_int(v);
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, true));  // row_shr:2
    x = __float_as_int(v);
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, true));  // row_shr:4
    x = __float_as_int(v);
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, true));  // row_shr:8
    x = __float_as_int(v);
    // rows 1 and 3 add lane 15 of rows 0 and 2; rows 2 and 3 then add lane 31
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false));  // row_bcast:15
    x = __float_as_int(v);
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false));  // row_bcast:31
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// Eight sums over the 64 lanes of a wavefront at once, by a butterfly that
// halves the data at every step: the halves swap (lanes 0-31 go on with
// values 0-3, lanes 32-63 with 4-7), then the rows of 16 (value 2r' + j in
// row r of pair j), then the 8-lane halves of each row, so that value q sums
// in lanes 8q..8q+7; three DPP steps finish each group.  26 instructions for
// the eight sums instead of seven per sum.  On return f[q] holds the total of
// value q in every lane.
__device__ __forceinline__ void wave_sum8(float (&f)[8])
{
    float s[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        // lanes 32-63 of f[j] trade places with lanes 0-31 of f[j + 4]
        const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(f[j]), __float_as_uint(f[j + 4]),
                                                        false, false);
        s[j] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
    float t[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        // odd rows of s[j] trade places with even rows of s[j + 2]
        const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(s[j]), __float_as_uint(s[j + 2]),
                                                        false, false);
        t[j] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
    const bool lo = (threadIdx.x & 15) < 8;
    const float x = lo ? t[1] : t[0], y = lo ? t[0] : t[1];
    float z = y + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x128, 0xf, 0xf, false));  // row_ror:8
    z += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(z), 0x141, 0xf, 0xf, false));  // row_half_mirror
    z += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(z), 0x4e, 0xf, 0xf, false));   // quad_perm 2,3,0,1
    z += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(z), 0xb1, 0xf, 0xf, false));   // quad_perm 1,0,3,2
#pragma unroll
    for (int q = 0; q < 8; ++q) f[q] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(z), 8 * q));
}

// Keys cubic-convolution weights, a = -0.5 (ba_interp2.cpp:204-213).
template <typename T>
__device__ __forceinline__ void keys_w(T t, T w[4])
{
    const T t2 = t * t, t3 = t2 * t;
    w[0] = T(0.5) * (-t + T(2) * t2 - t3);
    w[1] = T(0.5) * (T(2) - T(5) * t2 + T(3) * t3);
    w[2] = T(0.5) * (t + T(4) * t2 - T(3) * t3);
    w[3] = T(0.5) * (-t2 + t3);
}

template <typename T>
__device__ __forceinline__ void keys_dw(T t, T w[4])
{
    const T t2 = t * t;
    w[0] = T(0.5) * (T(-1) + T(4) * t - T(3) * t2);
    w[1] = T(0.5) * (T(-10) * t + T(9) * t2);
    w[2] = T(0.5) * (T(1) + T(8) * t - T(9) * t2);
    w[3] = T(0.5) * (T(-2) * t + T(3) * t2);
}

// Twice the Keys weights, 2 w(t), as explicit FMAs (the contraction is
// pinned, so every kernel rounds them alike).  Scaling by 2 is exact, so a
// gather with these weights is exactly 4 times the gather with w(t).
__device__ __forceinline__ void keys_w2(float t, float w[4])
{
    const float t2 = t * t, t3 = t2 * t;
    w[0] = fmaf(2.f, t2, -t) - t3;           // -t + 2 t^2 - t^3
    w[1] = fmaf(3.f, t3, fmaf(-5.f, t2, 2.f));  // 2 - 5 t^2 + 3 t^3
    w[2] = fmaf(-3.f, t3, fmaf(4.f, t2, t));    // t + 4 t^2 - 3 t^3
    w[3] = t3 - t2;                          // t^3 - t^2
}

// Keys gather of window offset (dx, dy) from a tile of row pitch tnx (the
// taps of the 4 x 4 support, weights from the split coordinates), at four
// times its value: the taps are combined with 2 w in both directions.
__device__ __forceinline__ float split_gather4(const SplitWarp &s, const float *tile, int tnx,
                                               float dx, float dy)
{
    const float su = fmaf(s.p1, dx, fmaf(s.p3, dy, s.fu));
    const float sv = fmaf(s.p2, dx, fmaf(s.p4, dy, s.fv));
    const float fsu = floorf(su), fsv = floorf(sv);
    float wu[4], wv[4];
    keys_w2(__builtin_amdgcn_fractf(su), wu);
    keys_w2(__builtin_amdgcn_fractf(sv), wv);
    const float iuf = (dx + s.cu) + fsu, ivf = (dy + s.cv) + fsv;
    const float *b = tile + (int)fmaf(ivf, (float)tnx, iuf);
    float g = 0.f;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const float *c4 = b + q * tnx;
        g = fmaf(wv[q], fmaf(wu[0], c4[0], fmaf(wu[1], c4[1], fmaf(wu[2], c4[2], wu[3] * c4[3]))), g);
    }
    return g;
}

// The same gather at its value (the 1/4 is exact).
__device__ __forceinline__ float split_gather(const SplitWarp &s, const float *tile, int tnx,
                                              float dx, float dy)
{
    return 0.25f * split_gather4(s, tile, tnx, dx, dy);
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// MATLAB mldivide for a small symmetric system: Cholesky when the diagonal is
// positive and the factorisation succeeds, otherwise LU with partial pivoting.
// A is n x n column-major; b and x have n entries.
template <int N>
__device__ inline void solve_small(const double *A_in, const double *b, double *x)
{
    double A[N * N], L[N * N];
    for (int i = 0; i < N * N; ++i) A[i] = A_in[i];
    bool spd = true;
    for (int i = 0; i < N; ++i)
        if (!(A[i + i * N] > 0)) spd = false;
    if (spd) {
        for (int i = 0; i < N * N; ++i) L[i] = 0.0;
        for (int j = 0; j < N && spd; ++j) {
            double d = A[j + j * N];
            for (int k = 0; k < j; ++k) d -= L[j + k * N] * L[j + k * N];
            if (!(d > 0)) {
                spd = false;
                break;
            }
            const double ljj = sqrt(d);
            L[j + j * N] = ljj;
            for (int i = j + 1; i < N; ++i) {
                double s = A[i + j * N];
                for (int k = 0; k < j; ++k) s -= L[i + k * N] * L[j + k * N];
                L[i + j * N] = s / ljj;
            }
        }
        if (spd) {
            double y[N];
            for (int i = 0; i < N; ++i) {
                double s = b[i];
                for (int k = 0; k < i; ++k) s -= L[i + k * N] * y[k];
                y[i] = s / L[i + i * N];
            }
            for (int i = N - 1; i >= 0; --i) {
                double s = y[i];
                for (int k = i + 1; k < N; ++k) s -= L[k + i * N] * x[k];
                x[i] = s / L[i + i * N];
            }
            return;
        }
    }
    double bb[N];
    for (int i = 0; i < N; ++i) bb[i] = b[i];
    for (int k = 0; k < N; ++k) {
        int p = k;
        double best = fabs(A[k + k * N]);
        for (int i = k + 1; i < N; ++i)
            if (fabs(A[i + k * N]) > best) {
                best = fabs(A[i + k * N]);
                p = i;
            }
        if (p != k) {
            for (int j = 0; j < N; ++j) {
                const double t = A[k + j * N];
                A[k + j * N] = A[p + j * N];
                A[p + j * N] = t;
            }
            const double t = bb[k];
            bb[k] = bb[p];
            bb[p] = t;
        }
        for (int i = k + 1; i < N; ++i) {
            const double l = A[i + k * N] / A[k + k * N];
            A[i + k * N] = l;
            for (int j = k + 1; j < N; ++j) A[i + j * N] -= l * A[k + j * N];
            bb[i] -= l * bb[k];
        }
    }
    for (int i = N - 1; i >= 0; --i) {
        double s = bb[i];
        for (int j = i + 1; j < N; ++j) s -= A[i + j * N] * x[j];
        x[i] = s / A[i + i * N];
    }
}

// The same factorisation kept in LDS: A (N x N, column-major) is overwritten
// by its Cholesky factor (kind 0) or by LU factors with row pivots (kind 1).
// Called by one thread; the system matrix is constant over the iterations, so
// it is factored once per subset and each iteration only substitutes.
template <int N>
__device__ inline void factor_lds(double *A, double *work, int *piv, int *kind)
{
    // work (N x N, LDS) receives the trial Cholesky factor so that A stays
    // intact for the LU fallback
    bool spd = true;
    for (int i = 0; i < N; ++i)
        if (!(A[i + i * N] > 0)) spd = false;
    if (spd) {
        for (int j = 0; j < N && spd; ++j) {
            double d = A[j + j * N];
            for (int k = 0; k < j; ++k) d -= work[j + k * N] * work[j + k * N];
            if (!(d > 0)) {
                spd = false;
                break;
            }
            const double ljj = sqrt(d);
            work[j + j * N] = ljj;
            for (int i = j + 1; i < N; ++i) {
                double s = A[i + j * N];
                for (int k = 0; k < j; ++k) s -= work[i + k * N] * work[j + k * N];
                work[i + j * N] = s / ljj;
            }
        }
        if (spd) {
            for (int j = 0; j < N; ++j)
                for (int i = j; i < N; ++i) A[i + j * N] = work[i + j * N];
            *kind = 0;
            return;
        }
    }
    for (int i = 0; i < N; ++i) piv[i] = i;
    for (int k = 0; k < N; ++k) {
        int p = k;
        double best = fabs(A[k + k * N]);
        for (int i = k + 1; i < N; ++i)
            if (fabs(A[i + k * N]) > best) {
                best = fabs(A[i + k * N]);
                p = i;
            }
        if (p != k) {
            for (int j = 0; j < N; ++j) {
                const double t = A[k + j * N];
                A[k + j * N] = A[p + j * N];
                A[p + j * N] = t;
            }
            const int t = piv[k];
            piv[k] = piv[p];
            piv[p] = t;
        }
        for (int i = k + 1; i < N; ++i) {
            const double l = A[i + k * N] / A[k + k * N];
            A[i + k * N] = l;
            for (int j = k + 1; j < N; ++j) A[i + j * N] -= l * A[k + j * N];
        }
    }
    *kind = 1;
}

template <int N>
__device__ inline void solve_factored(const double *A, const int *piv, int kind, const double *b,
                                      double *x)
{
    double y[N];
    if (kind == 0) {
        for (int i = 0; i < N; ++i) {
            double s = b[i];
            for (int k = 0; k < i; ++k) s -= A[i + k * N] * y[k];
            y[i] = s / A[i + i * N];
        }
        for (int i = N - 1; i >= 0; --i) {
            double s = y[i];
            for (int k = i + 1; k < N; ++k) s -= A[k + i * N] * x[k];
            x[i] = s / A[i + i * N];
        }
    } else {
        for (int i = 0; i < N; ++i) {
            double s = b[piv[i]];
            for (int k = 0; k < i; ++k) s -= A[i + k * N] * y[k];
            y[i] = s;
        }
        for (int i = N - 1; i >= 0; --i) {
            double s = y[i];
            for (int k = i + 1; k < N; ++k) s -= A[i + k * N] * x[k];
            x[i] = s / A[i + i * N];
        }
    }
}

// Statistics of one iteration from the residual-form sums
// acc = [sum r, sum r q, sum J_q r] over the window, r = g - kf f - ck and
// q = g + kf f - c1k (ck, c1k: (1 -/+ kf) cf rounded to FP32), and the
// set-up's sums about cf (F1c = sum (f - cf), F2c = sum (f - cf)^2,
// FJc = sum J (f - cf)): dg = gbar - cf, var = sum (g - gbar)^2 and
// T_q = sum J_q (g - cf), all in FP64.  With a = g - cf, e = f - cf:
// r = a - kf e - corr, q = a + kf e - corr1, so sum a, sum a^2 and
// sum J a follow exactly from the FP32 sums of the small r and r q.
template <int NJ>
__device__ __forceinline__ void residual_stats(double n, const float *acc, float kf, float ck,
                                               float c1k, float cf, double F1c, double F2c,
                                               const double *FJc, const double *SJ, double &dg,
                                               double &var, double (&T)[NJ])
{
    const double k = kf;
    const double corr = (double)ck - (1.0 - k) * (double)cf;
    const double corr1 = (double)c1k - (1.0 + k) * (double)cf;
    const double S1 = (double)acc[0] + k * F1c + n * corr;
    const double S2 = (double)acc[1] + k * k * F2c + corr1 * (S1 - k * F1c) +
                      corr * (S1 + k * F1c) - n * corr * corr1;
    dg = S1 / n;
    var = S2 - S1 * dg;
#pragma unroll
    for (int q = 0; q < NJ; ++q) T[q] = (double)acc[2 + q] + k * FJc[q] + corr * SJ[q];
}

// Inverse-compositional update W(P) <- W(P) o W(dP)^-1 with the closed-form
// inverse of funICGNQuadtree.m:393-411.  Returns false if det(dP) == 0.
__device__ inline bool ic_compose(double P[6], const double d[6])
{
    const double det = (1 + d[0]) * (1 + d[3]) - d[1] * d[2];
    if (!(det != 0)) return false;
    const double id = 1.0 / det;
    const double q1 = (-d[0] - d[0] * d[3] + d[1] * d[2]) * id;
    const double q2 = -d[1] * id;
    const double q3 = -d[2] * id;
    const double q4 = (-d[3] - d[0] * d[3] + d[1] * d[2]) * id;
    const double q5 = (-d[4] - d[3] * d[4] + d[2] * d[5]) * id;
    const double q6 = (-d[5] - d[0] * d[5] + d[1] * d[4]) * id;
    const double a11 = 1 + P[0], a12 = P[2], a13 = P[4];
    const double a21 = P[1], a22 = 1 + P[3], a23 = P[5];
    const double b11 = 1 + q1, b12 = q3, b13 = q5;
    const double b21 = q2, b22 = 1 + q4, b23 = q6;
    P[0] = (a11 * b11 + a12 * b21) - 1;
    P[1] = a21 * b11 + a22 * b21;
    P[2] = a11 * b12 + a12 * b22;
    P[3] = (a21 * b12 + a22 * b22) - 1;
    P[4] = a11 * b13 + a12 * b23 + a13;
    P[5] = a21 * b13 + a22 * b23 + a23;
    return true;
}

// Sum K values over the block.  Every thread passes its partials in v[];
// the totals land in out[] of thread 0 only.  red must hold K * (BS/64).
template <int BS, int K, typename T>
__device__ __forceinline__ void block_sum(T (&v)[K], T *red, T (&out)[K])
{
    // the wavefront's index in a scalar register (alike in every lane)
    const int lane = threadIdx.x & (kWave - 1), wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
#pragma unroll
    for (int k = 0; k < K; ++k) {
        T s;
        if constexpr (sizeof(T) == 4)
            s = wave_sum_dpp(v[k]);
        else
            s = wave_sum_dpp_f64(v[k]);
        if (lane == 0) red[wid * K + k] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            T s = red[k];
#pragma unroll
            for (int w = 1; w < BS / kWave; ++w) s += red[w * K + k];
            out[k] = s;
        }
    }
}

}  // namespace staq
