// Device helpers shared by the wavefront-per-subset IC-GN kernels
// (icgn_wave.hip, icgn_lean.hip): wavefront votes and sums, the sums over the
// wavefronts of one subset, the 6x6 Cholesky factor and solve, and the 2x2
// damped system of the 2-DOF step.  Both kernels use these same routines, so
// their common arithmetic is bit-identical.
#pragma once

#include "kernels.h"

namespace staq {
namespace {

template <typename T>
__device__ __forceinline__ T wave_allsum(T v)
{
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

__device__ __forceinline__ bool wave_any(bool p) { return __ballot(p) != 0ull; }

// In-register Cholesky of a 6x6 SPD matrix given by its packed lower
// triangle (column-major: (0,0),(1,0)..(5,0),(1,1)..); false if not SPD.
__device__ __forceinline__ bool chol6(double (&h)[21])
{
#define PK(r, c) h[(c) * 6 - ((c) * ((c) - 1)) / 2 + ((r) - (c))]
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        double d = PK(j, j);
#pragma unroll
        for (int k = 0; k < j; ++k) d -= PK(j, k) * PK(j, k);
        if (!(d > 0)) return false;
        const double l = sqrt(d);
        PK(j, j) = l;
        const double il = 1.0 / l;
#pragma unroll
        for (int i = j + 1; i < 6; ++i) {
            double s = PK(i, j);
#pragma unroll
            for (int k = 0; k < j; ++k) s -= PK(i, k) * PK(j, k);
            PK(i, j) = s * il;
        }
    }
#undef PK
    return true;
}

// Solve H d = -b given the packed Cholesky factor Lp of H (column-major lower
// triangle) and Ld = 1 / diag(Lp): forward, then backward substitution.
__device__ __forceinline__ void solve_chol6(const double *Lp, const double *Ld, const double (&b)[6],
                                            double (&d)[6])
{
    double y[6];
#define LP(r, c) Lp[(c) * 6 - ((c) * ((c) - 1)) / 2 + ((r) - (c))]
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        double s = -b[i];
#pragma unroll
        for (int k = 0; k < i; ++k) s -= LP(i, k) * y[k];
        y[i] = s * Ld[i];
        __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int i = 5; i >= 0; --i) {
        double s = y[i];
#pragma unroll
        for (int k = i + 1; k < 6; ++k) s -= LP(k, i) * d[k];
        d[i] = s * Ld[i];
        __builtin_amdgcn_sched_barrier(0);
    }
#undef LP
}

// Subset-wide sums over the NW wavefronts of a workgroup: f and d hold each
// lane's partials (reduced within the wavefront by DPP; eight floats together
// by the lane-swap butterfly of wave_sum8), iv each wavefront's own totals.  Every wavefront adds the same partials in the same order, so
// all of them hold bit-identical totals afterwards; the buffers alternate, so
// one barrier per exchange suffices.
template <int NW, int NF, int ND, int NI, typename Lds>
__device__ __forceinline__ void group_reduce(Lds &L, int &par, float *f, double *d, int *iv)
{
    if constexpr (NF == 8) {
        wave_sum8(*reinterpret_cast<float(*)[8]>(f));
    } else {
#pragma unroll
        for (int q = 0; q < NF; ++q) f[q] = wave_sum_dpp(f[q]);
    }
#pragma unroll
    for (int q = 0; q < ND; ++q) d[q] = wave_sum_dpp_f64(d[q]);
    if constexpr (NW > 1) {
        // the wavefront's index in a scalar register (alike in every lane)
        const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        if ((threadIdx.x & 63) == 0) {
#pragma unroll
            for (int q = 0; q < NF; ++q) L.xf[par][w][q] = f[q];
#pragma unroll
            for (int q = 0; q < ND; ++q) L.xd[par][w][q] = d[q];
#pragma unroll
            for (int q = 0; q < NI; ++q) L.xi[par][w][q] = iv[q];
        }
        __syncthreads();
        if constexpr (ND == 0 && NF > 24) {
            // many floats (the lean kernel's set-up sums): lane q < NF adds
            // the wavefronts' partials of value q, ints in lanes 56..
            constexpr int XW = sizeof(L.xf[0][0]) / sizeof(float);
            static_assert(NF <= XW && NF <= 56 && NI <= 8, "exchange layout");
            const int ln = threadIdx.x & 63;
            float vf = 0.f;
            int vi = 0;
            if (ln < NF) {
                float t = L.xf[par][0][ln];
#pragma unroll
                for (int ww = 1; ww < NW; ++ww) t += L.xf[par][ww][ln];
                vf = t;
            } else if (ln >= 56 && ln < 56 + NI) {
                int t = L.xi[par][0][ln - 56];
#pragma unroll
                for (int ww = 1; ww < NW; ++ww) t += L.xi[par][ww][ln - 56];
                vi = t;
            }
#pragma unroll
            for (int q = 0; q < NF; ++q) f[q] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(vf), q));
#pragma unroll
            for (int q = 0; q < NI; ++q) iv[q] = __builtin_amdgcn_readlane(vi, 56 + q);
            par ^= 1;
            return;
        } else if constexpr (ND > 0 || NW > 4) {
            // many values (the set-up sums) or many wavefronts: lane q adds
            // the wavefronts' partials of value q (floats in lanes 0..,
            // doubles in lanes 32.., ints in lanes 56..), then each total is
            // read from its lane; no lane holds every partial at once
            static_assert(NF <= 24 && ND <= 16 && NI <= 8, "exchange layout");
            const int ln = threadIdx.x & 63;
            float vf = 0.f;
            double vd = 0.0;
            int vi = 0;
            if (ln < NF) {
                float t = L.xf[par][0][ln];
#pragma unroll
                for (int ww = 1; ww < NW; ++ww) t += L.xf[par][ww][ln];
                vf = t;
            } else if (ln >= 32 && ln < 32 + ND) {
                double t = L.xd[par][0][ln - 32];
#pragma unroll
                for (int ww = 1; ww < NW; ++ww) t += L.xd[par][ww][ln - 32];
                vd = t;
            } else if (ln >= 56 && ln < 56 + NI) {
                int t = L.xi[par][0][ln - 56];
#pragma unroll
                for (int ww = 1; ww < NW; ++ww) t += L.xi[par][ww][ln - 56];
                vi = t;
            }
#pragma unroll
            for (int q = 0; q < NF; ++q) f[q] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(vf), q));
#pragma unroll
            for (int q = 0; q < ND; ++q) d[q] = readlane_f64(vd, 32 + q);
#pragma unroll
            for (int q = 0; q < NI; ++q) iv[q] = __builtin_amdgcn_readlane(vi, 56 + q);
            par ^= 1;
            return;
        }
#pragma unroll
        for (int q = 0; q < NF; ++q) {
            float t = L.xf[par][0][q];
#pragma unroll
            for (int ww = 1; ww < NW; ++ww) t += L.xf[par][ww][q];
            f[q] = t;
        }
#pragma unroll
        for (int q = 0; q < ND; ++q) {
            double t = L.xd[par][0][q];
#pragma unroll
            for (int ww = 1; ww < NW; ++ww) t += L.xd[par][ww][q];
            d[q] = t;
        }
#pragma unroll
        for (int q = 0; q < NI; ++q) {
            int t = L.xi[par][0][q];
#pragma unroll
            for (int ww = 1; ww < NW; ++ww) t += L.xi[par][ww][q];
            iv[q] = t;
        }
        par ^= 1;
    }
}

// 2x2 damped system of funICGN_Subpb1.m:111,391-394, inverted by lane 0.
template <typename Lds>
__device__ __forceinline__ void set_h2(Lds &L, double h11, double h12, double h22, double bf,
                                       double mu)
{
    const double s2 = 2.0 / (bf * bf);
    const double H2a = h11 * s2 + mu, H2b = h12 * s2, H2d = h22 * s2 + mu;
    const double dl = 1e-3 * (H2a > H2d ? H2a : H2d);
    if (threadIdx.x == 0) {
        const double A[4] = {H2a + dl, H2b, H2b, H2d + dl};
        for (int c = 0; c < 2; ++c) {
            double e[2] = {0, 0}, xs[2];
            e[c] = 1.0;
            solve_small<2>(A, e, xs);
            L.H2i[0 + c * 2] = xs[0];
            L.H2i[1 + c * 2] = xs[1];
        }
    }
}

}  // namespace
}  // namespace staq
