// Correlation quality factors of the integer searches (compute_qFactor,
// funIntegerSearch.m:211-264, called per search point at :130-131 and per box
// of the multigrid search's last level at funIntegerSearchMg.m:170-173), for
// gfx950.
//
// For a box: the whole normxcorr2 map of C = f(box_f) against D = g(box_g),
// every placement of C over the zero-padded D ((n1 + d1 - 1) x (n2 + d2 - 1)
// placements), with normxcorr2's definition
//   num = sum(C.*Dw) - sum(Dw) sum(C)/mn,
//   den = sqrt(sum((C - mean C)^2)) sqrt(max(sum(Dw.^2) - sum(Dw)^2/mn, 0)),
//   cc  = num/den where den > sqrt(eps(max den)), else 0,
// then, with cc_min = cc - min(cc) and peak = max(cc_min):
//   pce = peak^2 / mean(cc_min.^2)         (peak to correlation energy)
//   ppe = 1 / sum_k p_k log(1/p_k)         (peak to entropy), p_k the
//         fractions of cc_min in 30 equal bins over [0, peak] (bin k holds
//         k w <= x < (k+1) w, w = peak/30, the last bin closed).
//
// One workgroup per box.  C and D sit in LDS as FP32 (the images are FP32 in
// HBM); a placement sums over the overlap of C with D only (the zero ring
// adds nothing), FP32 x FP32 products accumulated in FP64, as the integer
// search kernels do.  Boxes whose C and D do not fit in LDS read them from
// HBM instead (the same arithmetic).  The map goes to a scratch buffer
// (num and den, then cc), so the three passes (max den; cc with its min and
// max; energy and histogram) each read it once.
#include "kernels.h"

namespace staq {

namespace {

constexpr int kQBins = 30;
constexpr int kQThreads = 256;

struct QArgs {
    const float *f, *g;
    int H, W;
    const long *bf, *bg;  // 4 per box: x0 x1 y0 y1, 1-based, inclusive
    const long *map0;     // first placement of each box in map
    double *map;          // 2 doubles per placement
    double *q;            // per box: pce, ppe
    int *ok;              // per box: 1 computed, 0 constant template
};

__device__ __forceinline__ double block_max(double v, double *red)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = red[0];
    for (int w = 1; w < kQThreads / 64; ++w) r = fmax(r, red[w]);
    __syncthreads();
    return r;
}

__device__ __forceinline__ double block_min(double v, double *red)
{
    return -block_max(-v, red);
}

__device__ __forceinline__ double block_sum(double v, double *red)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = red[0];
    for (int w = 1; w < kQThreads / 64; ++w) r += red[w];
    __syncthreads();
    return r;
}

template <bool LDS>
__global__ __launch_bounds__(kQThreads) void qfactor_kernel(QArgs a, const long *boxes)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ double red[kQThreads / 64];
    __shared__ int hist[kQBins];
    const long b = boxes[blockIdx.x];
    const long fx0 = a.bf[4 * b], fy0 = a.bf[4 * b + 2];
    const long gx0 = a.bg[4 * b], gy0 = a.bg[4 * b + 2];
    const int n1 = (int)(a.bf[4 * b + 1] - fx0 + 1), n2 = (int)(a.bf[4 * b + 3] - fy0 + 1);
    const int d1 = (int)(a.bg[4 * b + 1] - gx0 + 1), d2 = (int)(a.bg[4 * b + 3] - gy0 + 1);
    const int P1 = n1 + d1 - 1, P2 = n2 + d2 - 1, NP = P1 * P2;
    const float *fT = a.f + (fx0 - 1) + (size_t)(fy0 - 1) * a.H;
    const float *gD = a.g + (gx0 - 1) + (size_t)(gy0 - 1) * a.H;
    float *T = reinterpret_cast<float *>(smem);
    float *D = T + n1 * n2;
    if constexpr (LDS) {
        for (int k = threadIdx.x; k < n1 * n2; k += kQThreads) T[k] = fT[k % n1 + (size_t)(k / n1) * a.H];
        for (int k = threadIdx.x; k < d1 * d2; k += kQThreads) D[k] = gD[k % d1 + (size_t)(k / d1) * a.H];
    }
    const long tld = LDS ? n1 : a.H, dld = LDS ? d1 : a.H;
    const float *Tp = LDS ? T : fT;
    const float *Dp = LDS ? D : gD;
    if (threadIdx.x < kQBins) hist[threadIdx.x] = 0;
    __syncthreads();

    // template sums; a constant template is normxcorr2's error
    double s = 0.0, ss = 0.0, lo = INFINITY, hi = -INFINITY;
    for (int k = threadIdx.x; k < n1 * n2; k += kQThreads) {
        const double v = Tp[k % n1 + (size_t)(k / n1) * tld];
        s += v;
        ss = fma(v, v, ss);
        lo = fmin(lo, v);
        hi = fmax(hi, v);
    }
    const double sT = block_sum(s, red), sTT = block_sum(ss, red);
    const double tlo = block_min(lo, red), thi = block_max(hi, red);
    if (!(thi > tlo)) {
        if (threadIdx.x == 0) {
            a.ok[b] = 0;
            a.q[2 * b] = a.q[2 * b + 1] = NAN;
        }
        return;
    }
    const double mn = (double)n1 * n2;
    const double denT = sqrt(fmax(sTT - sT * sT / mn, 0.0));

    // pass 1: num and den of every placement, and the largest den
    double *map = a.map + 2 * a.map0[b];
    double dmax = 0.0;
    for (int p = threadIdx.x; p < NP; p += kQThreads) {
        const int s1 = p % P1, s2 = p / P1;
        const int e1 = s1 - (n1 - 1), e2 = s2 - (n2 - 1);  // D index of template (0, 0)
        const int i0 = e1 < 0 ? -e1 : 0, i1 = (d1 - e1) < n1 ? d1 - e1 : n1;
        const int j0 = e2 < 0 ? -e2 : 0, j1 = (d2 - e2) < n2 ? d2 - e2 : n2;
        double sta = 0.0, sa = 0.0, saa = 0.0;
        for (int j = j0; j < j1; ++j) {
            const float *tc = Tp + (size_t)j * tld;
            const float *dc = Dp + (size_t)(e2 + j) * dld;
            for (int i = i0; i < i1; ++i) {
                const double dv = dc[e1 + i];
                sta = fma((double)tc[i], dv, sta);
                sa += dv;
                saa = fma(dv, dv, saa);
            }
        }
        const double num = sta - sa * sT / mn;
        const double den = denT * sqrt(fmax(saa - sa * sa / mn, 0.0));
        map[2 * p] = num;
        map[2 * p + 1] = den;
        dmax = fmax(dmax, den);
    }
    dmax = block_max(dmax, red);
    // normxcorr2's tolerance sqrt(eps(max den))
    const double tol = sqrt(nextafter(dmax, INFINITY) - dmax);

    // pass 2: the map, its min and max
    double cmin = INFINITY, cmax = -INFINITY;
    for (int p = threadIdx.x; p < NP; p += kQThreads) {
        const double den = map[2 * p + 1];
        const double c = den > tol ? map[2 * p] / den : 0.0;
        map[2 * p] = c;
        cmin = fmin(cmin, c);
        cmax = fmax(cmax, c);
    }
    cmin = block_min(cmin, red);
    cmax = block_max(cmax, red);
    const double peak = cmax - cmin;
    const double width = peak / kQBins;

    // pass 3: correlation energy and the histogram of cc - min
    double e = 0.0;
    for (int p = threadIdx.x; p < NP; p += kQThreads) {
        const double x = map[2 * p] - cmin;
        e = fma(x, x, e);
        int k = 0;
        if (width > 0.0) {
            k = (int)floor(x / width);
            k = k < 0 ? 0 : (k > kQBins - 1 ? kQBins - 1 : k);
            if (x < (double)k * width) --k;
            else if (k < kQBins - 1 && x >= (double)(k + 1) * width) ++k;
        }
        atomicAdd(&hist[k], 1);
    }
    e = block_sum(e, red);
    if (threadIdx.x == 0) {
        double ent = 0.0;
        for (int k = 0; k < kQBins; ++k) {
            if (hist[k] == 0) continue;
            const double pk = (double)hist[k] / NP;
            ent += pk * log(1.0 / pk);
        }
        a.q[2 * b] = peak * peak / (e / NP);
        a.q[2 * b + 1] = 1.0 / ent;
        a.ok[b] = 1;
    }
}

}  // namespace

size_t qfactor_lds_bytes(long n1, long n2, long d1, long d2)
{
    return (size_t)(n1 * n2 + d1 * d2) * sizeof(float);
}

hipError_t qfactor_launch(const FrameDev &fr, const long *bf, const long *bg, const long *map0,
                          double *map, double *q, int *ok, const long *boxes_lds, long n_lds,
                          size_t lds_bytes, const long *boxes_hbm, long n_hbm, hipStream_t s)
{
    QArgs a;
    a.f = fr.f;
    a.g = fr.g;
    a.H = fr.H;
    a.W = fr.W;
    a.bf = bf;
    a.bg = bg;
    a.map0 = map0;
    a.map = map;
    a.q = q;
    a.ok = ok;
    if (n_lds > 0) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&qfactor_kernel<true>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(qfactor_kernel<true>, dim3((unsigned)n_lds), dim3(kQThreads), lds_bytes, s, a,
                           boxes_lds);
    }
    if (n_hbm > 0)
        hipLaunchKernelGGL(qfactor_kernel<false>, dim3((unsigned)n_hbm), dim3(kQThreads), 0, s, a,
                           boxes_hbm);
    return hipGetLastError();
}

}  // namespace staq
