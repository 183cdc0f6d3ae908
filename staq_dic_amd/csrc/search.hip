// Whole-field integer search (funIntegerSearch.m:82-203) with findpeak
// (findpeak.m:1-68), one workgroup per search point, for gfx950.
//
// For a grid point the template C = f(ii:ii+w, jj:jj+w) and the search block
// D = g(ii-r1 : ii+w+r1, jj-r2 : jj+w+r2) are staged in LDS (FP32; the images
// are FP32 in HBM).  Every kept template placement s1, s2 in [-1, 2r+1]
// (funIntegerSearch.m:134) is scored with normxcorr2's definition on a
// zero-padded D:
//   num = sum(T.*Dw) - sum(Dw) sum(T)/mn,
//   den = sqrt(sum((T - mean T)^2)) sqrt(max(sum(Dw.^2) - sum(Dw)^2/mn, 0)),
//   C   = num/den if den > sqrt(eps(max den)) else 0.
// FP32 x FP32 products are exact in FP64 and are accumulated in FP64, so the
// map (and therefore the integer argmax) matches an FP64 evaluation up to
// summation order.  The argmax takes the first maximum in column-major order
// (max(f(:))); thread 0 then fits findpeak's 3x3 quadratic in closed form.
//
// Two kernels compute the map.  search_strip_kernel (the default): the box
// sums sum(Dw), sum(Dw.^2) of every placement come from running sums, first
// down each column of D, then along the placements of each row (normxcorr2
// itself takes its local sums from running sums), and sum(T.*Dw) is a
// register-blocked correlation: a thread owns P consecutive placements down
// one column of placements and slides a register window of D past each
// template column, so every LDS read of T or D feeds several FP64
// multiply-adds.  search_kernel: one placement per thread, all three sums
// accumulated per pixel; it takes the geometries whose padded blocks exceed
// the strip kernel's LDS budget.
#include "kernels.h"

namespace staq {

namespace {

// Shared LDS of the map's final steps.
template <int BS>
struct FinishLds {
    double redd[BS / kWave * 3];
    double bestv[BS / kWave];
    int besti[BS / kWave];
    double tol;
    int best;
};

// Block-wide maximum of den (normxcorr2's tolerance sqrt(eps(max den))), the
// first maximum of the map in column-major order, then findpeak
// (findpeak.m:27-65) by thread 0 and the outputs of this search point.
template <int BS>
__device__ void search_finish(const SearchArgs &a, const double2 *map, int k1, int k2, int ii,
                              int jj, long o_idx, double dmax, FinishLds<BS> &F)
{
    constexpr int NW = BS / kWave;
    const int tid = threadIdx.x;
    double *redd = F.redd;
    double *bestv = F.bestv;
    int *besti = F.besti;
    double &sh_tol = F.tol;
    int &sh_best = F.best;
    // block max of den
    {
        double v = dmax;
#pragma unroll
        for (int o = kWave / 2; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
        if ((tid & (kWave - 1)) == 0) redd[tid / kWave] = v;
        __syncthreads();
        if (tid == 0) {
            double mx = redd[0];
            for (int w = 1; w < NW; ++w) mx = fmax(mx, redd[w]);
            sh_tol = sqrt(nextafter(mx, INFINITY) - mx);  // sqrt(eps(max den))
        }
        __syncthreads();
    }
    // first maximum in column-major order
    {
        const double tol = sh_tol;
        double bv = -INFINITY;
        int bi = 0x7fffffff;
        for (int q = tid; q < k1 * k2; q += BS) {
            const double2 nd = map[q];
            const double cval = nd.y > tol ? nd.x / nd.y : 0.0;
            if (cval > bv) {  // q increases, so ties keep the first index
                bv = cval;
                bi = q;
            }
        }
#pragma unroll
        for (int o = kWave / 2; o > 0; o >>= 1) {
            const double ov = __shfl_xor(bv, o, kWave);
            const int oi = __shfl_xor(bi, o, kWave);
            if (ov > bv || (ov == bv && oi < bi)) {
                bv = ov;
                bi = oi;
            }
        }
        if ((tid & (kWave - 1)) == 0) {
            bestv[tid / kWave] = bv;
            besti[tid / kWave] = bi;
        }
        __syncthreads();
        if (tid == 0) {
            double v = bestv[0];
            int i = besti[0];
            for (int w = 1; w < NW; ++w)
                if (bestv[w] > v || (bestv[w] == v && besti[w] < i)) {
                    v = bestv[w];
                    i = besti[w];
                }
            sh_best = i;
        }
        __syncthreads();
    }
    if (tid != 0) return;

    // findpeak.m:27-65 on the stored map
    const double tol = sh_tol;
    auto C = [&](int ka, int kb) {
        const double2 nd = map[ka + kb * k1];
        return nd.y > tol ? nd.x / nd.y : 0.0;
    };
    const int bi = sh_best;
    const int yp = bi % k1 + 1, xp = bi / k1 + 1;  // row (dim 1), column (dim 2)
    double xpeak = xp, ypeak = yp, maxf = C(yp - 1, xp - 1);
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
        // normal equations of the 3x3 quadratic fit in closed form
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
    a.u[o_idx] = ypeak - (double)(a.r1 + 2);  // zero_disp = ceil((2r+3)/2)
    a.v[o_idx] = xpeak - (double)(a.r2 + 2);
    a.ccmax[o_idx] = maxf;
    a.x0[o_idx] = (ii + ii + a.w) / 2.0;
    a.y0[o_idx] = (jj + jj + a.w) / 2.0;
}

}  // namespace

template <int BS>
__global__ __launch_bounds__(BS) void search_kernel(SearchArgs a)
{
    constexpr int NW = BS / kWave;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ double redd[NW * 3];
    __shared__ FinishLds<BS> F;
    __shared__ double sh_sT, sh_denT;
    __shared__ int sh_const;

    const int tid = threadIdx.x;
    const long t = a.point0 + blockIdx.x;
    const int ix = (int)(t % a.nx), iy = (int)(t / a.nx);
    const int ii = a.gx0 + ix * a.step, jj = a.gy0 + iy * a.step;  // 1-based corner
    const int m = a.w + 1;
    const int k1 = 2 * a.r1 + 3, k2 = 2 * a.r2 + 3;
    const int d1 = m + 2 * a.r1, d2 = m + 2 * a.r2;
    const int p1 = d1 + 2, p2 = d2 + 2;  // D with a one-pixel zero ring
    float *T = reinterpret_cast<float *>(smem);
    float *Dp = T + m * m;
    double2 *map = a.scratch + (size_t)blockIdx.x * (size_t)(k1 * k2);

    // stage template and padded search block
    for (int q = tid; q < m * m; q += BS) {
        const int i = q % m, j = q / m;
        T[q] = a.f[(ii + i - 1) + (size_t)(jj + j - 1) * a.H];
    }
    for (int q = tid; q < p1 * p2; q += BS) {
        const int i = q % p1, j = q / p1;
        const int di = i - 1, dj = j - 1;  // 0-based in D
        float v = 0.f;
        if (di >= 0 && di < d1 && dj >= 0 && dj < d2)
            v = a.g[(ii - a.r1 + di - 1) + (size_t)(jj - a.r2 + dj - 1) * a.H];
        Dp[q] = v;
    }
    __syncthreads();

    // template statistics (FP64)
    {
        double s[3] = {0, 0, 0};
        const float t0 = T[0];
        for (int q = tid; q < m * m; q += BS) {
            s[0] += T[q];
            s[2] += T[q] != t0 ? 1.0 : 0.0;
        }
        double o[3];
        block_sum<BS, 3>(s, redd, o);
        if (tid == 0) {
            sh_sT = o[0];
            sh_const = o[2] == 0.0;
        }
        __syncthreads();
        double s2[1] = {0};
        const double mT = sh_sT / (double)(m * m);
        for (int q = tid; q < m * m; q += BS) s2[0] += ((double)T[q] - mT) * ((double)T[q] - mT);
        double o2[1];
        block_sum<BS, 1>(s2, redd, o2);
        if (tid == 0) sh_denT = sqrt(o2[0]);
        __syncthreads();
    }
    const long o_idx = iy + (long)ix * a.ny;
    if (sh_const) {  // normxcorr2 rejects a constant template -> NaN (catch)
        if (tid == 0) {
            a.u[o_idx] = a.v[o_idx] = a.ccmax[o_idx] = NAN;
            a.x0[o_idx] = (ii + ii + a.w) / 2.0;
            a.y0[o_idx] = (jj + jj + a.w) / 2.0;
        }
        return;
    }

    // numerator / denominator for every kept placement
    const double mn = (double)(m * m), sT = sh_sT, denT = sh_denT;
    double dmax = 0.0;
    for (int q = tid; q < k1 * k2; q += BS) {
        const int ka = q % k1, kb = q / k1;  // placement s = k - 1, padded offset = k
        double sTA = 0.0, sA = 0.0, sAA = 0.0;
        for (int j = 0; j < m; ++j) {
            const float *Tc = T + j * m;
            const float *Dc = Dp + ka + (kb + j) * p1;
            for (int i = 0; i < m; ++i) {
                const double d = Dc[i];
                sTA = fma((double)Tc[i], d, sTA);
                sA += d;
                sAA = fma(d, d, sAA);
            }
        }
        const double num = sTA - sA * sT / mn;
        const double dA = fmax(sAA - sA * sA / mn, 0.0);
        const double den = denT * sqrt(dA);
        map[q] = make_double2(num, den);
        dmax = fmax(dmax, den);
    }
    search_finish<BS>(a, map, k1, k2, ii, jj, o_idx, dmax, F);
}

// Register-blocked strip kernel (see the file comment).  P placements per
// thread along dim 1; LDS holds T as MP x m (rows beyond m are zero) and D
// with its zero ring as PH x p2 (rows beyond p1 are zero), so every read of
// a thread's register window stays inside its column.
template <int BS, int P>
__global__ __launch_bounds__(BS) void search_strip_kernel(SearchArgs a)
{
    constexpr int NW = BS / kWave;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ double redd[NW * 3];
    __shared__ FinishLds<BS> F;
    __shared__ double sh_sT, sh_denT;
    __shared__ int sh_const;

    const int tid = threadIdx.x;
    const long t = a.point0 + blockIdx.x;
    const int ix = (int)(t % a.nx), iy = (int)(t / a.nx);
    const int ii = a.gx0 + ix * a.step, jj = a.gy0 + iy * a.step;  // 1-based corner
    const int m = a.w + 1;
    const int k1 = 2 * a.r1 + 3, k2 = 2 * a.r2 + 3;
    const int d1 = m + 2 * a.r1, d2 = m + 2 * a.r2;
    const int p1 = d1 + 2, p2 = d2 + 2;  // D with a one-pixel zero ring
    const int K1P = (k1 + P - 1) / P * P, MP = (m + P - 1) / P * P;
    const int PH = K1P + MP - 1;  // >= p1 = k1 + m - 1
    float *T = reinterpret_cast<float *>(smem);
    float *Dp = T + MP * m;
    double2 *map = a.scratch + (size_t)blockIdx.x * (size_t)(k1 * k2);
    double2 *cs = a.colsum + (size_t)blockIdx.x * (size_t)(k1 * p2);

    // stage the template (zero rows below m) and the padded search block
    for (int q = tid; q < MP * m; q += BS) {
        const int i = q % MP, j = q / MP;
        T[q] = i < m ? a.f[(ii + i - 1) + (size_t)(jj + j - 1) * a.H] : 0.f;
    }
    for (int q = tid; q < PH * p2; q += BS) {
        const int i = q % PH, j = q / PH;
        const int di = i - 1, dj = j - 1;  // 0-based in D
        float v = 0.f;
        if (di >= 0 && di < d1 && dj >= 0 && dj < d2)
            v = a.g[(ii - a.r1 + di - 1) + (size_t)(jj - a.r2 + dj - 1) * a.H];
        Dp[q] = v;
    }
    __syncthreads();

    // template statistics (FP64)
    {
        double s[3] = {0, 0, 0};
        const float t0 = T[0];
        for (int q = tid; q < m * m; q += BS) {
            const float v = T[(q % m) + (q / m) * MP];
            s[0] += v;
            s[2] += v != t0 ? 1.0 : 0.0;
        }
        double o[3];
        block_sum<BS, 3>(s, redd, o);
        if (tid == 0) {
            sh_sT = o[0];
            sh_const = o[2] == 0.0;
        }
        __syncthreads();
        double s2[1] = {0};
        const double mT = sh_sT / (double)(m * m);
        for (int q = tid; q < m * m; q += BS) {
            const double v = T[(q % m) + (q / m) * MP];
            s2[0] += (v - mT) * (v - mT);
        }
        double o2[1];
        block_sum<BS, 1>(s2, redd, o2);
        if (tid == 0) sh_denT = sqrt(o2[0]);
        __syncthreads();
    }
    const long o_idx = iy + (long)ix * a.ny;
    if (sh_const) {  // normxcorr2 rejects a constant template -> NaN (catch)
        if (tid == 0) {
            a.u[o_idx] = a.v[o_idx] = a.ccmax[o_idx] = NAN;
            a.x0[o_idx] = (ii + ii + a.w) / 2.0;
            a.y0[o_idx] = (jj + jj + a.w) / 2.0;
        }
        return;
    }

    // box sums, first down each column c of the padded block: for every
    // placement row ka, sum over rows ka .. ka+m-1 of D and D^2 (running sums)
    for (int c = tid; c < p2; c += BS) {
        const float *col = Dp + (size_t)c * PH;
        double s1 = 0.0, s2 = 0.0;
        for (int i = 0; i < m; ++i) {
            const double v = col[i];
            s1 += v;
            s2 = fma(v, v, s2);
        }
        cs[(size_t)c * k1] = make_double2(s1, s2);
        for (int ka = 1; ka < k1; ++ka) {
            const double vo = col[ka - 1], vn = col[ka + m - 1];
            s1 += vn - vo;
            s2 += fma(vn, vn, -vo * vo);
            cs[ka + (size_t)c * k1] = make_double2(s1, s2);
        }
    }
    __syncthreads();
    // ... then along each placement row: sum over columns kb .. kb+m-1
    for (int ka = tid; ka < k1; ka += BS) {
        double s1 = 0.0, s2 = 0.0;
        for (int j = 0; j < m; ++j) {
            const double2 v = cs[ka + (size_t)j * k1];
            s1 += v.x;
            s2 += v.y;
        }
        map[ka] = make_double2(s1, s2);
        for (int kb = 1; kb < k2; ++kb) {
            const double2 vo = cs[ka + (size_t)(kb - 1) * k1], vn = cs[ka + (size_t)(kb + m - 1) * k1];
            s1 += vn.x - vo.x;
            s2 += vn.y - vo.y;
            map[ka + (size_t)kb * k1] = make_double2(s1, s2);
        }
    }
    __syncthreads();

    // sum(T.*Dw) for strips of P placements; consecutive threads take
    // consecutive placement columns (LDS columns p2 apart: no bank conflict)
    const double mn = (double)(m * m), sT = sh_sT, denT = sh_denT;
    double dmax = 0.0;
    const int nstrip = (K1P / P) * k2;
    for (int st = tid; st < nstrip; st += BS) {
        const int kb = st % k2, ka0 = (st / k2) * P;
        double acc[P];
#pragma unroll
        for (int q = 0; q < P; ++q) acc[q] = 0.0;
        for (int j = 0; j < m; ++j) {
            const float *Dc = Dp + (size_t)(kb + j) * PH + ka0;
            const float *Tc = T + (size_t)j * MP;
            for (int i0 = 0; i0 < MP; i0 += P) {
                double d[2 * P - 1];
#pragma unroll
                for (int e = 0; e < 2 * P - 1; ++e) d[e] = Dc[i0 + e];
#pragma unroll
                for (int u = 0; u < P; ++u) {
                    const double tv = Tc[i0 + u];
#pragma unroll
                    for (int q = 0; q < P; ++q) acc[q] = fma(tv, d[u + q], acc[q]);
                }
            }
        }
#pragma unroll
        for (int q = 0; q < P; ++q) {
            const int ka = ka0 + q;
            if (ka < k1) {
                const double2 bs = map[ka + (size_t)kb * k1];
                const double num = acc[q] - bs.x * sT / mn;
                const double dA = fmax(bs.y - bs.x * bs.x / mn, 0.0);
                const double den = denT * sqrt(dA);
                map[ka + (size_t)kb * k1] = make_double2(num, den);
                dmax = fmax(dmax, den);
            }
        }
    }
    __syncthreads();
    search_finish<BS>(a, map, k1, k2, ii, jj, o_idx, dmax, F);
}

constexpr int kStripP = 8;

size_t search_strip_lds_bytes(int w, int r1, int r2)
{
    const size_t m = (size_t)w + 1, P = kStripP;
    const size_t k1 = 2 * (size_t)r1 + 3;
    const size_t K1P = (k1 + P - 1) / P * P, MP = (m + P - 1) / P * P;
    return sizeof(float) * (MP * m + (K1P + MP - 1) * (m + 2 * (size_t)r2 + 2));
}

size_t search_lds_bytes(int w, int r1, int r2)
{
    const size_t m = (size_t)w + 1;
    return sizeof(float) * (m * m + (m + 2 * r1 + 2) * (m + 2 * r2 + 2));
}

bool search_uses_strip(int w, int r1, int r2)
{
    return search_strip_lds_bytes(w, r1, r2) <= 150 * 1024;
}

hipError_t search_launch(const SearchArgs &a, long npoints, hipStream_t s)
{
    if (npoints <= 0) return hipSuccess;
    static bool attr = false;
    if (!attr) {
        hipFuncSetAttribute((const void *)search_kernel<256>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 8192);
        hipFuncSetAttribute((const void *)search_strip_kernel<256, kStripP>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 8192);
        hipFuncSetAttribute((const void *)search_strip_kernel<512, kStripP>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 8192);
        attr = true;
    }
    if (search_uses_strip(a.w, a.r1, a.r2)) {
        // a block that needs more than half the LDS runs alone on its CU:
        // eight wavefronts then keep two per SIMD to hide the FP64 latency
        const size_t lds = search_strip_lds_bytes(a.w, a.r1, a.r2);
        if (lds > 64 * 1024)
            hipLaunchKernelGGL((search_strip_kernel<512, kStripP>), dim3((unsigned)npoints),
                               dim3(512), lds, s, a);
        else
            hipLaunchKernelGGL((search_strip_kernel<256, kStripP>), dim3((unsigned)npoints),
                               dim3(256), lds, s, a);
    } else {
        const size_t lds = search_lds_bytes(a.w, a.r1, a.r2);
        hipLaunchKernelGGL(search_kernel<256>, dim3((unsigned)npoints), dim3(256), lds, s, a);
    }
    return hipGetLastError();
}

}  // namespace staq
