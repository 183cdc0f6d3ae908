// The multigrid integer search (funIntegerSearchMg.m:3-285) on the device:
// the two correlation kernels it needs.  The box bookkeeping between levels
// and the final gridfit stay with the caller, as in the reference.
//
//   * Level 1 (:17-27): normxcorr2 of the whole region of interest of f
//     against the same region of g, every placement of the template over the
//     zero-padded image (the "full" map, (n1+a1-1) x (n2+a2-1)), and findpeak
//     on it.  sum(T.*A) over every placement is a correlation, taken with
//     FP64 FFTs (hipFFT); the local sums of A and A.^2 are running sums, as
//     normxcorr2 takes them; the map is normalised, its first maximum in
//     column-major order found, and findpeak's 3x3 fit applied.
//   * Levels >= 2 (:151-170): for every box of a level, normxcorr2 of
//     C = f(box) against D = g(box +- pad), cropped to the placements with
//     shifts -pad-1 .. pad+1 in dim 1 and -pad-1 .. pad+2 in dim 2 (the
//     crop of :165, which keeps one column more on the far side), then
//     findpeak.  Boxes are cut into 64 x 64 template tiles; a workgroup per
//     tile sums T.*Dw, Dw, Dw.^2 for every kept placement over its tile (the
//     zero ring of D included), a second kernel adds the tiles of each box in
//     a fixed order and finishes the box.
#include <hipfft/hipfft.h>

#include "kernels.h"

namespace staq {

namespace {

// findpeak (findpeak.m:27-65) on a column-major map given by an accessor;
// (bi) is the first maximum in column-major order.  Returns xpeak (dim 2,
// column), ypeak (dim 1, row), max_f, all 1-based as in the reference.
template <typename Acc>
__device__ void findpeak_at(Acc C, long k1, long k2, long bi, double *out)
{
    const long yp = bi % k1 + 1, xp = bi / k1 + 1;
    double xpeak = (double)xp, ypeak = (double)yp, maxf = C(yp - 1, xp - 1);
    if (!(xp == 1 || xp == k2 || yp == 1 || yp == k1)) {
        double su = 0, sxu = 0, syu = 0, sxyu = 0, sxxu = 0, syyu = 0;
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy) {
                const double uu = C(yp - 1 + dy, xp - 1 + dx);
                su += uu;
                sxu += dx * uu;
                syu += dy * uu;
                sxyu += dx * dy * uu;
                sxxu += dx * dx * uu;
                syyu += dy * dy * uu;
            }
        const double A1 = (5.0 * su - 3.0 * sxxu - 3.0 * syyu) / 9.0;
        const double A2 = sxu / 6.0, A3 = syu / 6.0, A4 = sxyu / 4.0;
        const double A5 = -su / 3.0 + sxxu / 2.0, A6 = -su / 3.0 + syyu / 2.0;
        const double den = A4 * A4 - 4.0 * A5 * A6;
        double xo = (-A3 * A4 + 2.0 * A6 * A2) / den;
        double yo = -1.0 / den * (A4 * A2 - 2.0 * A5 * A3);
        if (!(fabs(xo) > 1 || fabs(yo) > 1)) {
            xo = round(1000.0 * xo) / 1000.0;
            yo = round(1000.0 * yo) / 1000.0;
            xpeak = xp + xo;
            ypeak = yp + yo;
            maxf = A1 + A2 * xo + A3 * yo + A4 * xo * yo + A5 * xo * xo + A6 * yo * yo;
        }
    }
    out[0] = xpeak;
    out[1] = ypeak;
    out[2] = maxf;
}

// ------------------------------------------------------------------ level 1

// rot180(T) and A into zero-padded complex FP64 arrays P1 x P2 (column-major)
__global__ void full_stage_kernel(const float *f, const float *g, int H, long fx0, long fy0,
                                  long n1, long n2, long gx0, long gy0, long a1, long a2, long P1,
                                  long P2, double2 *ft, double2 *fa)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= P1 * P2) return;
    const long i = q % P1, j = q / P1;
    double t = 0.0, v = 0.0;
    if (i < n1 && j < n2)
        t = f[(fx0 + (n1 - 1 - i) - 1) + (size_t)(fy0 + (n2 - 1 - j) - 1) * H];
    if (i < a1 && j < a2) v = g[(gx0 + i - 1) + (size_t)(gy0 + j - 1) * H];
    ft[q] = make_double2(t, 0.0);
    fa[q] = make_double2(v, 0.0);
}

__global__ void cmul_kernel(double2 *a, const double2 *b, long n)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const double2 x = a[q], y = b[q];
    a[q] = make_double2(x.x * y.x - x.y * y.y, x.x * y.y + x.y * y.x);
}

// Sums of A and A.^2 over the template-sized window ending at each map
// position (zero padding outside A): first down each column (one thread per
// column of A), then along each map row (one thread per row).
__global__ void full_colsum_kernel(const float *g, int H, long gx0, long gy0, long a1, long a2,
                                   long n1, long k1, double2 *cs)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a2) return;
    const float *col = g + (gx0 - 1) + (size_t)(gy0 + c - 1) * H;
    double s1 = 0.0, s2 = 0.0;
    for (long k = 0; k < k1; ++k) {  // window rows k-n1+1 .. k of A
        if (k < a1) {
            const double v = col[k];
            s1 += v;
            s2 = fma(v, v, s2);
        }
        const long o = k - n1;
        if (o >= 0 && o < a1) {
            const double v = col[o];
            s1 -= v;
            s2 -= v * v;
        }
        cs[k + c * k1] = make_double2(s1, s2);
    }
}

__global__ void full_rowsum_kernel(const double2 *cs, long k1, long k2, long a2, long n2,
                                   double2 *ls)
{
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= k1) return;
    double s1 = 0.0, s2 = 0.0;
    for (long l = 0; l < k2; ++l) {  // window columns l-n2+1 .. l of A
        if (l < a2) {
            const double2 v = cs[k + l * k1];
            s1 += v.x;
            s2 += v.y;
        }
        const long o = l - n2;
        if (o >= 0 && o < a2) {
            const double2 v = cs[k + o * k1];
            s1 -= v.x;
            s2 -= v.y;
        }
        ls[k + l * k1] = make_double2(s1, s2);
    }
}

// numerator and denominator in place: ls <- (num, den); block maxima of den
__global__ void full_numden_kernel(const double2 *xc, long P1, double2 *ls, long k1, long k2,
                                   double sT, double mn, double denT, double nfft, double *dmax)
{
    __shared__ double red[256];
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double d = 0.0;
    if (q < k1 * k2) {
        const long k = q % k1, l = q / k1;
        const double xcorr = xc[k + l * P1].x / nfft;  // unnormalised inverse FFT
        const double2 s = ls[q];
        const double num = xcorr - s.x * sT / mn;
        const double dA = fmax(s.y - s.x * s.x / mn, 0.0);
        d = denT * sqrt(dA);
        ls[q] = make_double2(num, d);
    }
    red[threadIdx.x] = d;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) dmax[blockIdx.x] = red[0];
}

__device__ __forceinline__ double ncc_at(const double2 *nd, long q, double tol)
{
    const double2 v = nd[q];
    return v.y > tol ? v.x / v.y : 0.0;
}

// first maximum in column-major order: per-block (value, index), then one block
// largest of n values by one workgroup (the block maxima of den)
__global__ void max_reduce_kernel(const double *v, long n, double *out)
{
    __shared__ double red[1024];
    double m = 0.0;
    for (long q = threadIdx.x; q < n; q += blockDim.x) m = fmax(m, v[q]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

__global__ void full_argmax_kernel(const double2 *nd, long n, const double *dmax_all,
                                   double *bv_out, long *bi_out)
{
    __shared__ double sv[256];
    __shared__ long si[256];
    const double mx = dmax_all[0];
    const double tol = sqrt(nextafter(mx, INFINITY) - mx);  // sqrt(eps(max den))
    double bv = -INFINITY;
    long bi = 0x7fffffffffffffffL;
    const long per = (n + gridDim.x - 1) / gridDim.x;
    const long q0 = (long)blockIdx.x * per, q1 = q0 + per < n ? q0 + per : n;
    for (long q = q0 + threadIdx.x; q < q1; q += blockDim.x) {
        const double c = ncc_at(nd, q, tol);
        if (c > bv) {
            bv = c;
            bi = q;
        }
    }
    sv[threadIdx.x] = bv;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            const double ov = sv[threadIdx.x + o];
            const long oi = si[threadIdx.x + o];
            if (ov > sv[threadIdx.x] || (ov == sv[threadIdx.x] && oi < si[threadIdx.x])) {
                sv[threadIdx.x] = ov;
                si[threadIdx.x] = oi;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        bv_out[blockIdx.x] = sv[0];
        bi_out[blockIdx.x] = si[0];
    }
}

__global__ void full_peak_kernel(const double2 *nd, long k1, long k2, const double *dmax_all,
                                 const double *bv, const long *bi, long nb, double *out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double mx = dmax_all[0];
    const double tol = sqrt(nextafter(mx, INFINITY) - mx);
    double v = bv[0];
    long i = bi[0];
    for (long b = 1; b < nb; ++b)
        if (bv[b] > v || (bv[b] == v && bi[b] < i)) {
            v = bv[b];
            i = bi[b];
        }
    auto C = [&](long r, long c) { return ncc_at(nd, r + c * k1, tol); };
    findpeak_at(C, k1, k2, i, out);
}

// template statistics of the region (sum T, sum T^2, count of values other
// than the first): per-block partials, then one block adds them in order
constexpr int kStatBlocks = 512;

__global__ void tstat_kernel(const float *f, int H, long fx0, long fy0, long n1, long n2,
                             double *part)
{
    __shared__ double r0[256], r1[256], r2[256];
    const long n = n1 * n2;
    const float t0 = f[(fx0 - 1) + (size_t)(fy0 - 1) * H];
    double s = 0.0, ss = 0.0, dif = 0.0;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (long)gridDim.x * blockDim.x) {
        const long i = q % n1, j = q / n1;
        const double v = f[(fx0 + i - 1) + (size_t)(fy0 + j - 1) * H];
        s += v;
        ss = fma(v, v, ss);
        dif += v != t0 ? 1.0 : 0.0;
    }
    r0[threadIdx.x] = s;
    r1[threadIdx.x] = ss;
    r2[threadIdx.x] = dif;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            r0[threadIdx.x] += r0[threadIdx.x + o];
            r1[threadIdx.x] += r1[threadIdx.x + o];
            r2[threadIdx.x] += r2[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x] = r0[0];
        part[3 * blockIdx.x + 1] = r1[0];
        part[3 * blockIdx.x + 2] = r2[0];
    }
}

__global__ void tstat_sum_kernel(const double *part, int nb, double *acc)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double a = 0.0, b = 0.0, c = 0.0;
    for (int k = 0; k < nb; ++k) {
        a += part[3 * k];
        b += part[3 * k + 1];
        c += part[3 * k + 2];
    }
    acc[0] = a;
    acc[1] = b;
    acc[2] = c;
}

// ------------------------------------------------------------- levels >= 2

constexpr int kTile = 64;

struct BoxArgs {
    const float *f, *g;
    int H, W;
    int pad;                 // D = g(box_g +- pad)
    int kp1, kp2;            // kept placements: 2 pad + 3 by 2 pad + 4
    const long *bf, *bg;     // boxes, 4 per box: x0 x1 y0 y1 (1-based, inclusive)
    const int *tile0;        // first tile of each box (prefix sums), nbox + 1 entries
    long nbox;
    double *part;            // per tile: 3 sums per placement, then sum T, sum T^2, min, max
    double *out;             // per box: xpeak, ypeak, max_f
    int *ok;                 // per box: 1 computed, 0 error (the reference's catch: NaN)
};

// One workgroup per template tile: T tile and its D support in LDS, one thread
// per kept placement summing over the tile's pixels.
__global__ __launch_bounds__(256) void box_tile_kernel(BoxArgs a)
{
    __shared__ float T[kTile * kTile];
    __shared__ float D[(kTile + 16) * (kTile + 16)];
    __shared__ long box_sh;
    const long t = blockIdx.x;
    if (threadIdx.x == 0) {  // box of this tile (tile0 is increasing)
        long lo = 0, hi = a.nbox - 1;
        while (lo < hi) {
            const long mid = (lo + hi + 1) / 2;
            if (a.tile0[mid] <= t) lo = mid; else hi = mid - 1;
        }
        box_sh = lo;
    }
    __syncthreads();
    const long b = box_sh;
    const long fx0 = a.bf[4 * b], fx1 = a.bf[4 * b + 1], fy0 = a.bf[4 * b + 2], fy1 = a.bf[4 * b + 3];
    const long gx0 = a.bg[4 * b] - a.pad, gy0 = a.bg[4 * b + 2] - a.pad;
    const long n1 = fx1 - fx0 + 1, n2 = fy1 - fy0 + 1;
    const long d1 = n1 + 2 * a.pad, d2 = n2 + 2 * a.pad;
    const long nt1 = (n1 + kTile - 1) / kTile;
    const long lt = t - a.tile0[b];
    const long ti = (lt % nt1) * kTile, tj = (lt / nt1) * kTile;  // tile origin in T
    const int m1 = (int)(n1 - ti < kTile ? n1 - ti : kTile), m2 = (int)(n2 - tj < kTile ? n2 - tj : kTile);
    const int s1 = m1 + a.kp1 - 1, s2 = m2 + a.kp2 - 1;  // D support of the tile
    for (int q = threadIdx.x; q < kTile * kTile; q += blockDim.x) {
        const int i = q % kTile, j = q / kTile;
        T[q] = (i < m1 && j < m2) ? a.f[(fx0 + ti + i - 1) + (size_t)(fy0 + tj + j - 1) * a.H] : 0.f;
    }
    // placement (k, l) puts the template's top-left at D(k - 1, l - 1) (0-based,
    // -1 is the zero ring); the tile's pixel (i, j) meets D(ti + i + k - 1, tj + j + l - 1)
    for (int q = threadIdx.x; q < s1 * s2; q += blockDim.x) {
        const int i = q % s1, j = q / s1;
        const long di = ti + i - 1, dj = tj + j - 1;
        float v = 0.f;
        if (di >= 0 && di < d1 && dj >= 0 && dj < d2)
            v = a.g[(gx0 + di - 1) + (size_t)(gy0 + dj - 1) * a.H];
        D[i + j * (kTile + 16)] = v;
    }
    __syncthreads();
    const int nk = a.kp1 * a.kp2;
    double *part = a.part + (size_t)t * (3 * nk + 4);
    for (int q = threadIdx.x; q < nk; q += blockDim.x) {
        const int k = q % a.kp1, l = q / a.kp1;
        double sta = 0.0, sa = 0.0, saa = 0.0;
        for (int j = 0; j < m2; ++j) {
            const float *Tc = T + j * kTile;
            const float *Dc = D + k + (j + l) * (kTile + 16);
            for (int i = 0; i < m1; ++i) {
                const double d = Dc[i];
                sta = fma((double)Tc[i], d, sta);
                sa += d;
                saa = fma(d, d, saa);
            }
        }
        part[3 * q] = sta;
        part[3 * q + 1] = sa;
        part[3 * q + 2] = saa;
    }
    if (threadIdx.x == 0) {
        double s = 0.0, ss = 0.0, mn = INFINITY, mx = -INFINITY;
        for (int j = 0; j < m2; ++j)
            for (int i = 0; i < m1; ++i) {
                const double v = T[i + j * kTile];
                s += v;
                ss = fma(v, v, ss);
                mn = fmin(mn, v);
                mx = fmax(mx, v);
            }
        part[3 * nk] = s;
        part[3 * nk + 1] = ss;
        part[3 * nk + 2] = mn;
        part[3 * nk + 3] = mx;
    }
}

// One workgroup per box: add its tiles in order, normalise, argmax, findpeak.
__global__ __launch_bounds__(256) void box_finish_kernel(BoxArgs a)
{
    __shared__ double nd[2 * 256];
    __shared__ double dmax_s[256];
    __shared__ double st[4];
    const long b = blockIdx.x;
    const int nk = a.kp1 * a.kp2;
    const long t0 = a.tile0[b], t1 = a.tile0[b + 1];
    const long n1 = a.bf[4 * b + 1] - a.bf[4 * b] + 1, n2 = a.bf[4 * b + 3] - a.bf[4 * b + 2] + 1;
    if (threadIdx.x == 0) {
        double s = 0.0, ss = 0.0, mn = INFINITY, mx = -INFINITY;
        for (long t = t0; t < t1; ++t) {
            const double *p = a.part + (size_t)t * (3 * nk + 4) + 3 * nk;
            s += p[0];
            ss += p[1];
            mn = fmin(mn, p[2]);
            mx = fmax(mx, p[3]);
        }
        st[0] = s;
        st[1] = ss;
        st[2] = mn;
        st[3] = mx;
    }
    __syncthreads();
    const double mn = (double)(n1 * n2), sT = st[0];
    const double denT = sqrt(fmax(st[1] - sT * sT / mn, 0.0));
    double dm = 0.0;
    for (int q = threadIdx.x; q < nk; q += blockDim.x) {
        double sta = 0.0, sa = 0.0, saa = 0.0;
        for (long t = t0; t < t1; ++t) {
            const double *p = a.part + (size_t)t * (3 * nk + 4) + 3 * q;
            sta += p[0];
            sa += p[1];
            saa += p[2];
        }
        const double num = sta - sa * sT / mn;
        const double dA = fmax(saa - sa * sa / mn, 0.0);
        const double den = denT * sqrt(dA);
        nd[2 * q] = num;
        nd[2 * q + 1] = den;
        dm = fmax(dm, den);
    }
    dmax_s[threadIdx.x] = dm;
    __syncthreads();
    if (threadIdx.x != 0) return;
    double mx = 0.0;
    for (int i = 0; i < (int)blockDim.x; ++i) mx = fmax(mx, dmax_s[i]);
    if (!(st[3] > st[2])) {  // normxcorr2 rejects a constant template
        a.ok[b] = 0;
        return;
    }
    const double tol = sqrt(nextafter(mx, INFINITY) - mx);
    auto C = [&](long r, long c) {
        const long q = r + c * a.kp1;
        return nd[2 * q + 1] > tol ? nd[2 * q] / nd[2 * q + 1] : 0.0;
    };
    double bv = -INFINITY;
    long bi = 0;
    for (int q = 0; q < nk; ++q) {
        const double c = C(q % a.kp1, q / a.kp1);
        if (c > bv) {
            bv = c;
            bi = q;
        }
    }
    findpeak_at(C, a.kp1, a.kp2, bi, a.out + 3 * b);
    a.ok[b] = 1;
}

long fft_size(long n)
{
    // smallest size >= n whose prime factors are 2, 3, 5 and 7
    for (long s = n;; ++s) {
        long r = s;
        for (long p : {2L, 3L, 5L, 7L})
            while (r % p == 0) r /= p;
        if (r == 1) return s;
    }
}

}  // namespace

void xcorr_full_sizes(long n1, long n2, long a1, long a2, size_t bytes[4])
{
    const long k1 = n1 + a1 - 1, k2 = n2 + a2 - 1;
    const long P1 = fft_size(k1), P2 = fft_size(k2);
    const long nb_d = (k1 * k2 + 255) / 256;
    bytes[0] = (size_t)P1 * P2 * 16;                            // FFT of rot180(T), then the product
    bytes[1] = (size_t)P1 * P2 * 16;                            // FFT of A
    bytes[2] = (size_t)k1 * (size_t)(k2 > a2 ? k2 : a2) * 32;  // column sums, then local sums
    bytes[3] = (size_t)(nb_d + 2 * 1024 + 16 + 3 * kStatBlocks) * 8;  // reductions
}

int xcorr_full(const FrameDev &fr, const long box_f[4], const long box_g[4], void *const buf[4],
               hipStream_t s, double out[3], char *err, size_t errlen)
{
    const long H = fr.H;
    const long n1 = box_f[1] - box_f[0] + 1, n2 = box_f[3] - box_f[2] + 1;
    const long a1 = box_g[1] - box_g[0] + 1, a2 = box_g[3] - box_g[2] + 1;
    const long k1 = n1 + a1 - 1, k2 = n2 + a2 - 1;
    const long P1 = fft_size(k1), P2 = fft_size(k2);
    const long nb_d = (k1 * k2 + 255) / 256, nb_a = 1024;
    double2 *ft = static_cast<double2 *>(buf[0]), *fa = static_cast<double2 *>(buf[1]);
    double2 *cs = static_cast<double2 *>(buf[2]), *ls = cs + (size_t)k1 * (k2 > a2 ? k2 : a2);
    double *red = static_cast<double *>(buf[3]);
    double *dmax = red, *bv = red + nb_d, *stt = red + nb_d + 2 * nb_a;
    long *bi = reinterpret_cast<long *>(red + nb_d + nb_a);
    double *dmax_all = stt + 8, *spart = stt + 16;
    auto hip_err = [&](hipError_t e, const char *what) {
        snprintf(err, errlen, "%s: %s", what, hipGetErrorString(e));
        return -5;
    };

    const long n = P1 * P2;
    hipLaunchKernelGGL(full_stage_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, fr.f,
                       fr.g, (int)H, box_f[0], box_f[2], n1, n2, box_g[0], box_g[2], a1, a2, P1,
                       P2, ft, fa);
    hipLaunchKernelGGL(tstat_kernel, dim3(kStatBlocks), dim3(256), 0, s, fr.f, (int)H, box_f[0],
                       box_f[2], n1, n2, spart);
    hipLaunchKernelGGL(tstat_sum_kernel, dim3(1), dim3(64), 0, s, spart, kStatBlocks, stt);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_err(e, "xcorr stage");
    hipfftHandle plan;
    hipfftResult r = hipfftPlan2d(&plan, (int)P2, (int)P1, HIPFFT_Z2Z);  // dim 2 is the slow one
    if (r != HIPFFT_SUCCESS) {
        snprintf(err, errlen, "hipfftPlan2d: hipFFT error %d", (int)r);
        return -5;
    }
    hipfftSetStream(plan, s);
    auto z = [](double2 *p) { return reinterpret_cast<hipfftDoubleComplex *>(p); };
    r = hipfftExecZ2Z(plan, z(ft), z(ft), HIPFFT_FORWARD);
    if (r == HIPFFT_SUCCESS) r = hipfftExecZ2Z(plan, z(fa), z(fa), HIPFFT_FORWARD);
    if (r == HIPFFT_SUCCESS) {
        hipLaunchKernelGGL(cmul_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ft, fa, n);
        r = hipfftExecZ2Z(plan, z(ft), z(ft), HIPFFT_BACKWARD);
    }
    hipfftDestroy(plan);
    if (r != HIPFFT_SUCCESS) {
        snprintf(err, errlen, "hipfftExecZ2Z: hipFFT error %d", (int)r);
        return -5;
    }
    hipLaunchKernelGGL(full_colsum_kernel, dim3((unsigned)((a2 + 255) / 256)), dim3(256), 0, s, fr.g,
                       (int)H, box_g[0], box_g[2], a1, a2, n1, k1, cs);
    hipLaunchKernelGGL(full_rowsum_kernel, dim3((unsigned)((k1 + 255) / 256)), dim3(256), 0, s, cs,
                       k1, k2, a2, n2, ls);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_err(e, "xcorr local sums");
    double st[3];
    e = hipMemcpyAsync(st, stt, 24, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_err(e, "xcorr template statistics");
    if (st[2] == 0.0) {  // constant template: normxcorr2 errors (the caller's catch)
        out[0] = out[1] = out[2] = NAN;
        return 0;
    }
    const double mn = (double)(n1 * n2);
    const double denT = sqrt(fmax(st[1] - st[0] * st[0] / mn, 0.0));
    hipLaunchKernelGGL(full_numden_kernel, dim3((unsigned)nb_d), dim3(256), 0, s, ft, P1, ls, k1, k2,
                       st[0], mn, denT, (double)n, dmax);
    hipLaunchKernelGGL(max_reduce_kernel, dim3(1), dim3(1024), 0, s, dmax, nb_d, dmax_all);
    hipLaunchKernelGGL(full_argmax_kernel, dim3((unsigned)nb_a), dim3(256), 0, s, ls, k1 * k2,
                       dmax_all, bv, bi);
    hipLaunchKernelGGL(full_peak_kernel, dim3(1), dim3(64), 0, s, ls, k1, k2, dmax_all, bv, bi,
                       nb_a, stt);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, stt, 24, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_err(e, "xcorr peak");
    return 0;
}

size_t xcorr_box_part_doubles(int pad) { return (size_t)(3 * (2 * pad + 3) * (2 * pad + 4) + 4); }

hipError_t xcorr_boxes_launch(const FrameDev &fr, int pad, long nb, long ntile, const long *bf,
                              const long *bg, const int *tile0, double *part, double *out, int *ok,
                              hipStream_t s)
{
    BoxArgs a;
    a.f = fr.f;
    a.g = fr.g;
    a.H = fr.H;
    a.W = fr.W;
    a.pad = pad;
    a.kp1 = 2 * pad + 3;
    a.kp2 = 2 * pad + 4;
    a.bf = bf;
    a.bg = bg;
    a.tile0 = tile0;
    a.nbox = nb;
    a.part = part;
    a.out = out;
    a.ok = ok;
    hipLaunchKernelGGL(box_tile_kernel, dim3((unsigned)ntile), dim3(256), 0, s, a);
    hipLaunchKernelGGL(box_finish_kernel, dim3((unsigned)nb), dim3(256), 0, s, a);
    return hipGetLastError();
}

int xcorr_box_tiles(long n1, long n2) { return (int)(((n1 + kTile - 1) / kTile) * ((n2 + kTile - 1) / kTile)); }

}  // namespace staq
