// Batched inverse-compositional Gauss-Newton (IC-GN) subset fitting for gfx950.
//
//   MODE 6: funICGNQuadtree.m:1-451 (6-DOF affine, ZNSSD) and, with plain = 1,
//           the uniform-grid twin func/funICGN.m:1-323.
//   MODE 2: funICGN_Subpb1.m:1-474 (translation-only step of the augmented-
//           Lagrangian loop with penalty mu, affine part fixed at FOld).
//
// One workgroup fits one subset; a launch covers a whole batch (all quadtree
// nodes of one ADMM iteration, possibly across frame pairs).  With a work
// counter the grid is persistent and pulls subsets from a device queue.
//
// Per subset:
//   setup   - load the (w+1)^2 window of f*mask, the masked-pixel count
//             decides 102 / window growth; 4-connected region of the mask
//             through the seed (bwselect); FP64 sums for H (21 entries), the
//             f statistics and the constant parts of the steepest-descent
//             vector.
//   iterate - corner bounds check in FP64 (the warp is affine so the
//             extremes of u22/v22 are at the window corners), stage the
//             bounding tile of g in LDS with coalesced loads, Keys bicubic
//             from LDS, and ONE fused FP32 pass giving count, sum(g-c),
//             sum((g-c)^2) and sum(J_k (g-c)); thread 0 turns those into
//             b = bottomf * sum(J r) in FP64, runs the norm test, solves
//             H dP = -b (FP64 Cholesky, LU fallback) and composes the warp.
//   masked  - subsets that touch the mask (or whose g has exact zeros) take
//             the exact path: bwselect flood fills in LDS and two-pass FP64
//             statistics, following the reference line by line.
#include <type_traits>

#include "kernels.h"
#include "flood.h"

namespace staq {



struct Ctl {
    double P[6];
    double x0, y0;
    double Hm[36];     // factor of H (6-DOF) / scratch (2-DOF)
    double Lw[36];     // trial Cholesky factor
    int piv[6], fkind;
    double H21[21];
    double meanf, bottomf, sabs;
    double SJ[6], FJ[6];
    double normNew, normAbs, normInit;
    double tH[4];      // 2-DOF damped system
    double bvec[6];    // steepest-descent vector of the current iteration
    double gmean, bottomg, ng, nf;
    double FJc[6], F1c, F2c;  // sums about cf (residual form of the FP32 sums)
    double kap;               // bottomg / bottomf of the last iteration
    SplitWarp sw;             // FP32 warp coordinates of this iteration
    float cf;
    int rform;                // the window has no masked pixel: residual form of S1, S2
    int exact;                // a stopping test fell inside the guard band
    int tile_u0, tile_v0, tnx, tny, use_tile;
    int tile_ok;       // the LDS tile holds g over (tile_u0, tile_v0, tnx, tny)
    int reload;        // this iteration's support left the resident tile
    int xa, ya, wx, wy, nx, ny, npix, w0;
    int count, it, state, status, sum_bw2;
    long idx;
    int slot;
    int path;          // 0 fused, 1 masked (flood fill + exact), 2 exact without flood fill
    int fsel;          // 2-DOF: f, grad f, H come from the last g region (funICGN_Subpb1.m:171-190)
    int nanP;
    int sel0_all;
    int all_sel;
    int lit;           // the exact path's statistics are degenerate: literal per-pixel passes
    long item;
};

__device__ __forceinline__ bool win_ok(const FrameDev &fr, double x0, double y0, int wx, int wy)
{
    const double xa = x0 - wx * 0.5, xb = x0 + wx * 0.5, ya = y0 - wy * 0.5, yb = y0 + wy * 0.5;
    if (xa != floor(xa) || ya != floor(ya) || xb != floor(xb) || yb != floor(yb)) return false;
    return xa >= 4 && ya >= 4 && xb <= fr.H - 3 && yb <= fr.W - 3;
}

// Row/column of pixel p in a column-major window with nx rows.
__device__ __forceinline__ void pix_ij(int p, int nx, float rcp, int &i, int &j)
{
    j = (int)(((float)p + 0.5f) * rcp);
    i = p - j * nx;
}

// Connected component of bw through (seed_r, seed_c) (1-based, row = dim 1),
// 4- or 8-connected, by iterated propagation in LDS.
template <int BS>
__device__ void flood_fill(const unsigned char *bw, unsigned char *out, int nr, int nc,
                           int seed_r, int seed_c, int conn, int npix, float rcp)
{
    if (nr <= 64 && nc <= 64) {
        // the first wavefront: column j of the window as one 64-bit word in
        // lane j, the bit-parallel fill, the words back to bytes
        if (threadIdx.x < 64) {
            const int lane = threadIdx.x;
            unsigned long long w = 0ull;
            if (lane < nc)
                for (int i = 0; i < nr; ++i)
                    w |= (unsigned long long)(bw[i + lane * nr] != 0) << i;
            const unsigned long long sel = wave_flood(w, nc, seed_r, seed_c, conn);
            if (lane < nc)
                for (int i = 0; i < nr; ++i) out[i + lane * nr] = (unsigned char)((sel >> i) & 1ull);
        }
        __syncthreads();
        return;
    }
    for (int p = threadIdx.x; p < npix; p += BS) out[p] = 0;
    __syncthreads();
    if (threadIdx.x == 0 && seed_r >= 1 && seed_r <= nr && seed_c >= 1 && seed_c <= nc) {
        const int s = (seed_r - 1) + (seed_c - 1) * nr;
        if (bw[s]) out[s] = 1;
    }
    __syncthreads();
    for (;;) {
        int changed = 0;
        for (int p = threadIdx.x; p < npix; p += BS) {
            if (!bw[p] || out[p]) continue;
            int i, j;
            pix_ij(p, nr, rcp, i, j);
            bool hit = false;
            for (int dj = -1; dj <= 1 && !hit; ++dj)
                for (int di = -1; di <= 1; ++di) {
                    if (!di && !dj) continue;
                    if (conn == 4 && di && dj) continue;
                    const int qi = i + di, qj = j + dj;
                    if (qi < 0 || qi >= nr || qj < 0 || qj >= nc) continue;
                    if (out[qi + qj * nr]) {
                        hit = true;
                        break;
                    }
                }
            if (hit) {
                out[p] = 1;
                changed = 1;
            }
        }
        if (!__syncthreads_or(changed)) break;
    }
}

// DfDx, DfDy at window pixel p (image offset gi): from the LDS stage when the
// workgroup holds one (gl), else from HBM
__device__ __forceinline__ float2 grad_px(bool gl, const float2 *gradf, const FrameDev &fr, int p, size_t gi)
{
    return gl ? gradf[p] : make_float2(gld(fr.dfdx, gi), gld(fr.dfdy, gi));
}

// EXACT: the FP64 re-solve of subsets whose stopping test fell inside the
// guard band (and of windows that outgrew their bin): the bicubic gather is
// evaluated in FP64 (from the FP32 image, as every path reads it) and every
// iteration takes the two-pass FP64 statistics and the per-pixel residual
// of funICGNQuadtree.m:269-371 / funICGN_Subpb1.m:270-379, so the iteration
// count is the one an FP64 evaluation gives.  The re-solve asks for two waves
// per SIMD (256 VGPRs at most): left alone the compiler takes 276 and one
// workgroup per CU, so that long re-solves queue behind each other.
template <int BS, int MODE, bool EXACT>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(EXACT ? 2 : 1))) void icgn_kernel(IcgnArgs a)
{
    constexpr int NW = BS / kWave;
    constexpr int NJ = MODE;  // steepest-descent components
    using GT = typename std::conditional<EXACT, double, float>::type;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ Ctl c;
    __shared__ double redd[NW * 21];
    __shared__ float redf[NW * 12];
    __shared__ long item_sh;

    GT *gbuf = reinterpret_cast<GT *>(smem);
    // the re-solve may stage the window's DfDx, DfDy after g (gl)
    const bool gl = EXACT && a.grad_lds;
    unsigned char *after_g = smem + sizeof(GT) * (size_t)a.npix_cap;
    float2 *gradf = reinterpret_cast<float2 *>(after_g);
    float *tile = reinterpret_cast<float *>(after_g + (gl ? sizeof(float2) * (size_t)a.npix_cap : 0));
    unsigned char *tm = reinterpret_cast<unsigned char *>(tile + a.tile_cap);
    unsigned char *sel0 = tm + a.npix_cap;
    unsigned char *sel = sel0 + a.npix_cap;
    unsigned char *bwb = sel + a.npix_cap;
    const int tid = threadIdx.x;
    // fallback launch with nothing deferred: leave before touching the queue
    if (a.defer_flag && *(volatile const int *)a.defer_flag == 0) return;

    for (long pass = 0;; ++pass) {
        // ---------------------------------------------------------- pick work
        if (tid == 0) {
            long it;
            if (a.work_counter)
                it = a.item0 + atomicAdd(a.work_counter, 1);
            else
                it = pass == 0 ? a.item0 + blockIdx.x : a.n_items + a.item0;
            item_sh = it;
        }
        __syncthreads();
        const long item = item_sh;
        if (item >= a.item0 + a.n_items) break;
        // the fallback launch takes the subsets the main launches listed
        if (a.defer_list && a.only_deferred && item >= *(volatile const int *)a.defer_flag) break;

        // ------------------------------------------------------------ setup
        if (tid == 0) {
            const long idx = (a.defer_list && a.only_deferred)
                                 ? (long)a.defer_list[a.list_back ? a.N - 1 - item : item]
                                 : (a.order ? a.order[item] : item);
            c.idx = idx;
            c.slot = a.slot ? a.slot[idx] : 0;
            double x0 = a.coords[idx], y0 = a.coords[idx + a.N];
            if (!a.plain) {
                x0 = round(x0);  // LocalICGNQuadtree.m:69
                y0 = round(y0);
            }
            c.x0 = x0;
            c.y0 = y0;
            const int wx = a.wx[idx];
            const int wy = MODE == 2 ? a.wy[idx] : wx;
            c.wx = wx;
            c.wy = wy;
            c.w0 = wx;
            if (MODE == 6) {
                c.P[0] = c.P[1] = c.P[2] = c.P[3] = 0.0;
                c.P[4] = a.U0[2 * idx];
                c.P[5] = a.U0[2 * idx + 1];
            } else {
                for (int k = 0; k < 4; ++k) c.P[k] = a.FOld[4 * idx + k];
                c.P[4] = a.UOld[2 * idx];
                c.P[5] = a.UOld[2 * idx + 1];
            }
            c.state = 0;
            c.status = 0;
            c.it = 0;
            c.normNew = 1;
            c.normAbs = 1;
            c.normInit = 0;
            c.sum_bw2 = 0;
            c.sel0_all = 1;
            c.fsel = 0;
            c.kap = 1.0;
            c.exact = 0;
            c.tile_ok = 0;
            c.rform = 0;
            const FrameDev fr = a.frames[c.slot];
            if (a.only_deferred && a.status[idx] != ST_DEFER && a.status[idx] != ST_EXACT) {
                c.state = 4;  // done by the main launch
            } else if (!win_ok(fr, x0, y0, wx, wy)) {
                c.status = ST_ERROR;
                c.state = 2;
            } else if ((wx + 1) * (wy + 1) > a.npix_cap) {
                c.status = a.defer ? ST_DEFER : ST_ERROR;
                c.state = a.defer ? 3 : 2;
            } else {
                c.nx = wx + 1;
                c.ny = wy + 1;
                c.npix = c.nx * c.ny;
                c.xa = (int)(x0 - wx / 2);
                c.ya = (int)(y0 - wy / 2);
            }
        }
        __syncthreads();
        const FrameDev fr = a.frames[c.slot];
        const int Hd = fr.H - 6;

        if (c.state == 0) {
            // masked-pixel count: zeros of f*mask (funICGNQuadtree.m:61,
            // funICGN_Subpb1.m:68) or of DfDx (funICGN.m:54)
            {
                const int nx = c.nx, npix = c.npix, xa = c.xa, ya = c.ya;
                const float rcp = 1.0f / nx;
                float cnt = 0.f;
                for (int p = tid; p < npix; p += BS) {
                    int i, j;
                    pix_ij(p, nx, rcp, i, j);
                    const int x = xa + i, y = ya + j;
                    if (MODE == 6 && a.plain)
                        cnt += gld(fr.dfdx, (size_t)((x - 4) + (y - 4) * Hd)) == 0.f;
                    else
                        cnt += fidx(fr.f, fr.H, x, y) * fidx(fr.mask, fr.H, x, y) == 0.f;
                }
                float v[1] = {cnt}, o[1];
                block_sum<BS, 1>(v, redf, o);
                if (tid == 0) {
                    const int count = (int)o[0];
                    c.count = count;
                    const double n0 = (double)c.nx * c.ny;
                    const double cap = (MODE == 6 && a.plain) ? 0.5 * n0 : 0.4 * n0;
                    if (!(count < cap)) {
                        c.status = ST_MASKED;
                        c.state = 2;
                    } else {
                        int wx = c.wx, wy = c.wy;
                        bool grow;
                        if (MODE == 6 && a.plain) {
                            grow = count > 0.1 * n0;
                            if (grow) wx = wy = (int)round(sqrt(2.0) * wx);
                        } else {
                            grow = count > 0;
                            if (grow) {
                                wx = 2 * (int)ceil(sqrt(n0 / (n0 - count)) * wx / 2.0);
                                if (MODE == 6) {
                                    wy = wx;
                                } else {
                                    const double n1 = (double)(wx + 1) * (wy + 1);
                                    wy = 2 * (int)ceil(sqrt(n1 / (n1 - count)) * wy / 2.0);
                                }
                            }
                        }
                        if (grow) {
                            if (!win_ok(fr, c.x0, c.y0, wx, wy)) {
                                c.status = ST_ERROR;
                                c.state = 2;
                            } else if ((wx + 1) * (wy + 1) > a.npix_cap) {
                                c.status = a.defer ? ST_DEFER : ST_ERROR;
                                c.state = a.defer ? 3 : 2;
                            } else {
                                c.wx = wx;
                                c.wy = wy;
                                c.nx = wx + 1;
                                c.ny = wy + 1;
                                c.npix = c.nx * c.ny;
                                c.xa = (int)(c.x0 - wx / 2);
                                c.ya = (int)(c.y0 - wy / 2);
                            }
                        }
                        c.sel0_all = (a.plain && MODE == 6) || count == 0;
                    }
                }
                __syncthreads();
            }
        }

        const int nx = c.nx, ny = c.ny, npix = c.npix, xa = c.xa, ya = c.ya;
        const float rcp = 1.0f / nx;
        const float hx = c.wx * 0.5f, hy = c.wy * 0.5f;
        // bwselect seed: (floor((wx+1)/2), floor((wy+1)/2)) passed as (c, r):
        // the column index is floor((wx+1)/2) and the row floor((wy+1)/2).
        const int seed_c = (c.wx + 1) / 2, seed_r = (c.wy + 1) / 2;

        if (c.state == 0 && !c.sel0_all) {
            // :83-88 — 4-connected region of the window mask through the seed
            for (int p = tid; p < npix; p += BS) {
                int i, j;
                pix_ij(p, nx, rcp, i, j);
                tm[p] = fidx(fr.mask, fr.H, xa + i, ya + j) != 0.f;
            }
            __syncthreads();
            flood_fill<BS>(tm, sel0, nx, ny, seed_r, seed_c, 4, npix, rcp);
        }

        if (gl && c.state == 0) {
            // DfDx, DfDy of the window, read once: the window keeps its size
            // over the iterations (a new size ends the subset), so every pass
            // below reads them from LDS; 4 loads in flight per thread
            for (int p0 = 0; p0 < npix; p0 += BS * 4) {
                float2 q4[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int p = p0 + tid + BS * q;
                    q4[q] = make_float2(0.f, 0.f);
                    if (p < npix) {
                        int i, j;
                        pix_ij(p, nx, rcp, i, j);
                        const size_t gi = (size_t)((xa + i - 4) + (ya + j - 4) * Hd);
                        q4[q] = make_float2(gld(fr.dfdx, gi), gld(fr.dfdy, gi));
                    }
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int p = p0 + tid + BS * q;
                    if (p < npix) gradf[p] = q4[q];
                }
            }
            __syncthreads();
        }

        // f-side sums in FP64 (:91-103, :119-120): statistics and the constant
        // parts of b first, the 21 Hessian entries in a second pass
        if (c.state == 0) {
            constexpr int K = (MODE == 6) ? 16 : 11;
            double v[K];
#pragma unroll
            for (int k = 0; k < K; ++k) v[k] = 0.0;
            const bool all = c.sel0_all;
            for (int p = tid; p < npix; p += BS) {
                int i, j;
                pix_ij(p, nx, rcp, i, j);
                const int x = xa + i, y = ya + j;
                const bool s = all || sel0[p];
                const size_t gi = (size_t)((x - 4) + (y - 4) * Hd);
                const float2 gq = s ? grad_px(gl, gradf, fr, p, gi) : make_float2(0.f, 0.f);
                const double fx = gq.x, fy = gq.y;
                double fv = fidx(fr.f, fr.H, x, y);
                if (!(MODE == 6 && a.plain)) fv *= (double)fidx(fr.mask, fr.H, x, y);
                if (!s) fv = 0.0;
                const bool nz = (MODE == 6 && a.plain) || fabs(fv) > 1e-10;
                v[0] += nz ? 1.0 : 0.0;
                v[1] += nz ? fv : 0.0;
                v[2] += nz ? fv * fv : 0.0;
                v[3] += fabs(fv);
                if (MODE == 6) {
                    const double dx = i - (double)hx, dy = j - (double)hy;
                    const double J[6] = {fx * dx, fy * dx, fx * dy, fy * dy, fx, fy};
#pragma unroll
                    for (int k = 0; k < 6; ++k) {
                        v[4 + k] += J[k];
                        v[10 + k] += J[k] * fv;
                    }
                } else {
                    v[4] += fx * fx;
                    v[5] += fx * fy;
                    v[6] += fy * fy;
                    v[7] += fx;
                    v[8] += fy;
                    v[9] += fx * fv;
                    v[10] += fy * fv;
                }
            }
            double o[K];
            block_sum<BS, K>(v, redd, o);
            if (tid == 0) {
                const double n = o[0], sf = o[1], sf2 = o[2];
                const double mf = n > 0 ? sf / n : NAN;
                const double bf = n >= 2 ? sqrt(fmax(sf2 - sf * sf / n, 0.0)) : (n == 1 ? 0.0 : NAN);
                c.meanf = mf;
                c.bottomf = bf;
                c.sabs = o[3];
                c.cf = isfinite(mf) ? (float)mf : 0.f;
                // residual form: no masked pixel and every pixel in f's nonzero set
                c.rform = c.sel0_all && c.count == 0 && n == (double)c.npix;
                const double cfd = c.cf;
                c.F1c = sf - n * cfd;
                c.F2c = sf2 - 2.0 * cfd * sf + n * cfd * cfd;
                if (MODE == 6) {
                    for (int k = 0; k < 6; ++k) {
                        c.SJ[k] = o[4 + k];
                        c.FJ[k] = o[10 + k] - mf * o[4 + k];
                        c.FJc[k] = o[10 + k] - cfd * o[4 + k];
                    }
                } else {
                    const double s2 = 2.0 / (bf * bf);
                    const double H2a = o[4] * s2 + a.mu, H2b = o[5] * s2, H2d = o[6] * s2 + a.mu;
                    const double dl = 1e-3 * (H2a > H2d ? H2a : H2d);
                    c.tH[0] = H2a + dl;
                    c.tH[1] = c.tH[2] = H2b;
                    c.tH[3] = H2d + dl;
                    c.SJ[0] = o[7];
                    c.SJ[1] = o[8];
                    c.FJ[0] = o[9] - mf * o[7];
                    c.FJ[1] = o[10] - mf * o[8];
                    c.FJc[0] = o[9] - cfd * o[7];
                    c.FJc[1] = o[10] - cfd * o[8];
                }
            }
            __syncthreads();
        }
        if (MODE == 6 && c.state == 0) {
            double v[21];
#pragma unroll
            for (int k = 0; k < 21; ++k) v[k] = 0.0;
            const bool all = c.sel0_all;
            for (int p = tid; p < npix; p += BS) {
                int i, j;
                pix_ij(p, nx, rcp, i, j);
                const int x = xa + i, y = ya + j;
                const bool s = all || sel0[p];
                const size_t gi = (size_t)((x - 4) + (y - 4) * Hd);
                const float2 gq = s ? grad_px(gl, gradf, fr, p, gi) : make_float2(0.f, 0.f);
                const double fx = gq.x, fy = gq.y;
                const double dx = i - (double)hx, dy = j - (double)hy;
                const double J[6] = {fx * dx, fy * dx, fx * dy, fy * dy, fx, fy};
                int t = 0;
#pragma unroll
                for (int cc = 0; cc < 6; ++cc)
#pragma unroll
                    for (int r = cc; r < 6; ++r) v[t++] += J[r] * J[cc];
            }
            double o[21];
            block_sum<BS, 21>(v, redd, o);
            if (tid == 0) {
                int t = 0;
                for (int cc = 0; cc < 6; ++cc)
                    for (int r = cc; r < 6; ++r) {
                        c.H21[t] = o[t];
                        c.Hm[r + cc * 6] = c.Hm[cc + r * 6] = o[t];
                        ++t;
                    }
                factor_lds<6>(c.Hm, c.Lw, c.piv, &c.fkind);
                if (a.guard2 > 0) {
                    // an ill-conditioned or indefinite H: the FP64 re-solve
                    double dg6[6];
                    for (int q = 0; q < 6; ++q) dg6[q] = c.Hm[q + q * 6];
                    if (c.fkind != 0 || pivots_spread(dg6)) {
                        c.exact = 1;
                        c.state = 3;
                    }
                }
            }
            __syncthreads();
        }

        // ---------------------------------------------------------- iterate
        while (c.state == 0) {
            if (tid == 0) {
                const double tol = a.tol;
                if (!(c.it <= a.max_it && c.normNew > tol && c.normAbs > tol)) {
                    c.state = 1;
                } else {
                    c.it += 1;
                    double *P = c.P;
                    if (MODE == 6 && !a.plain && c.it > 1 && c.count > 0) {
                        // funICGNQuadtree.m:139-145: the window is re-sized
                        // from the last g region without reloading f; a new
                        // size throws at :363 (status -1) unless the new grid
                        // leaves the image first (:154-157, status 101).
                        const double wn =
                            2.0 * ceil(sqrt((double)(c.w0 + 1) * (c.w0 + 1) / (double)c.sum_bw2) *
                                       c.wx / 2.0);
                        if (wn != (double)c.wx) {
                            if (!isfinite(wn) || c.x0 - wn / 2 < 1 || c.x0 + wn / 2 > fr.H ||
                                c.y0 - wn / 2 < 1 || c.y0 + wn / 2 > fr.W) {
                                c.status = ST_ERROR;
                                c.state = 2;
                            } else {
                                const double h = wn / 2;
                                bool out = false;
                                for (int q = 0; q < 4; ++q) {
                                    const double dx = (q & 1) ? h : -h, dy = (q & 2) ? h : -h;
                                    const double u = (1 + P[0]) * dx + P[2] * dy + (c.x0 + P[4]);
                                    const double w = P[1] * dx + (1 + P[3]) * dy + (c.y0 + P[5]);
                                    if (u < 3 || u > fr.H - 2 || w < 3 || w > fr.W - 2) out = true;
                                }
                                if (out) {
                                    c.normNew = 1e6;
                                    c.state = 1;
                                } else {
                                    c.status = ST_ERROR;
                                    c.state = 2;
                                }
                            }
                        }
                    }
                    if (c.state == 0) {
                        bool nanP = false;
                        for (int k = 0; k < 6; ++k) nanP |= isnan(P[k]);
                        if (nanP) {
                            // NaN coordinates pass the bounds test and give NaN
                            // g: logical(NaN) throws (quadtree / Subpb1), the
                            // plain solver's norm turns NaN.
                            if (MODE == 2 || !a.plain) {
                                c.status = ST_ERROR;
                                c.state = 2;
                            } else if (c.it == 1) {
                                c.normNew = 0;  // NaN initial norm -> relative 0
                                c.normAbs = NAN;
                                c.state = 1;
                            } else {
                                c.normNew = NAN;
                                c.normAbs = NAN;
                                c.state = 1;
                            }
                        }
                    }
                    if (c.state == 0) {
                        double umin = INFINITY, umax = -INFINITY, vmin = INFINITY, vmax = -INFINITY;
                        bool out = false, edge = false;
                        // the near-edge hand-over only where a guard band is set
                        // (not in the FP64 re-solve, which decides every test)
                        const bool guard = a.guard2 > 0;
                        for (int q = 0; q < 4; ++q) {
                            const double dx = (q & 1) ? hx : -hx, dy = (q & 2) ? hy : -hy;
                            const double u = (1 + P[0]) * dx + P[2] * dy + (c.x0 + P[4]);
                            const double w = P[1] * dx + (1 + P[3]) * dy + (c.y0 + P[5]);
                            if (u < 3 || u > fr.H - 2 || w < 3 || w > fr.W - 2) out = true;
                            if (guard)
                                edge |= near_edge(u, 3.0, fr.H - 2.0) | near_edge(w, 3.0, fr.W - 2.0);
                            umin = fmin(umin, u);
                            umax = fmax(umax, u);
                            vmin = fmin(vmin, w);
                            vmax = fmax(vmax, w);
                        }
                        if (edge) {
                            // the bounds test decided within FP32 reach: FP64 re-solve
                            c.exact = 1;
                            c.state = 3;
                        } else if (out) {
                            c.normNew = 1e6;
                            c.state = 1;
                        } else {
                            const int tu0 = (int)floor(umin) - 2, tv0 = (int)floor(vmin) - 2;
                            const int nxt = (int)floor(umax) - (int)floor(umin) + 6;
                            const int nyt = (int)floor(vmax) - (int)floor(vmin) + 6;
                            // the resident tile serves while it holds the new
                            // support with the same margins (the samples are the
                            // same, so the gather is unchanged)
                            const bool keep = c.tile_ok && tu0 >= c.tile_u0 && tv0 >= c.tile_v0 &&
                                              tu0 + nxt <= c.tile_u0 + c.tnx &&
                                              tv0 + nyt <= c.tile_v0 + c.tny;
                            if (!keep) {
                                c.tile_u0 = tu0;
                                c.tile_v0 = tv0;
                                c.tnx = nxt;
                                c.tny = nyt;
                                c.use_tile = nxt * nyt <= a.tile_cap;
                                c.tile_ok = c.use_tile;
                            }
                            c.reload = !keep && c.use_tile;
                            // FP32 warp coordinates for the FP32 gather (the
                            // re-solve gathers in FP64 from P itself)
                            if constexpr (!EXACT) c.sw = split_warp(P, c.x0, c.y0, c.tile_u0, c.tile_v0);
                        }
                    }
                }
            }
            __syncthreads();
            if (c.state != 0) break;

            const int tu0 = c.tile_u0, tv0 = c.tile_v0, tnx = c.tnx, tny = c.tny;
            const bool use_tile = c.use_tile;
            if (use_tile && c.reload) {
                // flat over the tile, 4 independent loads in flight per thread
                const int total = tnx * tny;
                const float rcpn = 1.0f / tnx;
                for (int t0 = 0; t0 < total; t0 += BS * 4) {
                    float v[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int t = t0 + tid + BS * q;
                        v[q] = 0.f;
                        if (t < total) {
                            const int sj = (int)(((float)t + 0.5f) * rcpn), ri = t - sj * tnx;
                            const int y = min(max(tv0 + sj, 1), fr.W), x = min(max(tu0 + ri, 1), fr.H);
                            v[q] = gld(fr.g, (size_t)((x - 1) + (y - 1) * fr.H));
                        }
                    }
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int t = t0 + tid + BS * q;
                        if (t < total) tile[t] = v[q];
                    }
                }
            }
            __syncthreads();

            // gather + fused sums
            if constexpr (EXACT) {
                // FP64 gather of every window pixel, fused with the sums of
                // this iteration about c = cf: n, sum (g - c), sum (g - c)^2
                // over the nonzero g, T_q = sum J_q (g - c) over the f side,
                // and the counts of zeros and NaNs.  A window without a zero
                // of g (and a 2-DOF f side that is not rebuilt) takes b from
                // these sums in one pass; the others take the passes of the
                // exact path below
                const double *P = c.P;
                const double cd = c.cf;
                const bool whole = MODE == 6 && a.plain;
                const bool all = c.sel0_all;
                // v[K], v[K + 1]: the counts of zeros and NaNs of g (one reduction)
                constexpr int K = 3 + NJ;
                double v[K + 2];
#pragma unroll
                for (int k = 0; k < K + 2; ++k) v[k] = 0.0;
                for (int p = tid; p < npix; p += BS) {
                    int i, j;
                    pix_ij(p, nx, rcp, i, j);
                    const double dx = i - (double)hx, dy = j - (double)hy;
                    const double u = (1 + P[0]) * dx + P[2] * dy + (c.x0 + P[4]);
                    const double vv = P[1] * dx + (1 + P[3]) * dy + (c.y0 + P[5]);
                    const double fu = floor(u), fv = floor(vv);
                    double wu[4], wv[4];
                    keys_w(u - fu, wu);
                    keys_w(vv - fv, wv);
                    const int iu = (int)fu - 1, iv = (int)fv - 1;  // 1-based first taps
                    double g = 0.0;
                    if (use_tile) {
                        // the same FP32 samples of g, staged in LDS (origin tu0, tv0)
                        const float *b = tile + (iu - tu0) + (iv - tv0) * tnx;
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            const float *cq = b + q * tnx;
                            double sq = 0.0;
#pragma unroll
                            for (int r = 0; r < 4; ++r) sq = fma(wu[r], (double)cq[r], sq);
                            g = fma(wv[q], sq, g);
                        }
                    } else {
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            const int y = min(max(iv + q, 1), fr.W);
                            const float *col = fr.g + (size_t)(y - 1) * fr.H;
                            double sq = 0.0;
#pragma unroll
                            for (int r = 0; r < 4; ++r)
                                sq = fma(wu[r], (double)gld(col, (size_t)(min(max(iu + r, 1), fr.H) - 1)), sq);
                            g = fma(wv[q], sq, g);
                        }
                    }
                    gbuf[p] = g;
                    v[K] += g == 0.0 ? 1.0 : 0.0;
                    v[K + 1] += isnan(g) ? 1.0 : 0.0;
                    const bool nz = whole || fabs(g) > 1e-10;
                    const double e = g - cd;
                    v[0] += nz ? 1.0 : 0.0;
                    v[1] += nz ? e : 0.0;
                    v[2] += nz ? e * e : 0.0;
                    const bool sf = all || sel0[p];
                    if (sf) {
                        const size_t gi = (size_t)((xa + i - 4) + (ya + j - 4) * Hd);
                        const float2 gq = grad_px(gl, gradf, fr, p, gi);
                        const double fx = gq.x, fy = gq.y;
                        if (MODE == 6) {
                            v[3] += fx * dx * e;
                            v[4] += fy * dx * e;
                            v[5] += fx * dy * e;
                            v[6] += fy * dy * e;
                            v[7] += fx * e;
                            v[8] += fy * e;
                        } else {
                            v[3] += fx * e;
                            v[4] += fy * e;
                        }
                    }
                }
                double od[K + 2];
                block_sum<BS, K + 2>(v, redd, od);
                if (tid == 0) {
                    const bool quad = !(MODE == 6 && a.plain);
                    if (quad && od[K + 1] > 0) {
                        c.status = ST_ERROR;  // logical(NaN) throws
                        c.state = 2;
                    } else if (quad && (c.count > 0 || od[K] > 0)) {
                        c.path = 1;
                    } else if (MODE == 2 && c.fsel) {
                        c.path = 2;
                    } else {
                        // b = bottomf sum J (fhat - ghat) = FJ - (bf/bg) (T - (gbar - c) SJ)
                        // (funICGNQuadtree.m:269-270, 362-371) in FP64
                        const double n = od[0], S1 = od[1], S2 = od[2];
                        const double dg = S1 / n, var = S2 - S1 * dg, bf = c.bottomf;
                        if (!(n >= 2) || !(bf > 0) || !(var > 0) || !isfinite(bf) ||
                            !isfinite(var) || !isfinite(dg)) {
                            c.path = 2;  // literal per-pixel residual for degenerate statistics
                        } else {
                            const double bg = sqrt(var);
                            const double ratio = bf / bg;
                            c.kap = bg / bf;
                            for (int k = 0; k < NJ; ++k)
                                c.bvec[k] = c.FJ[k] - ratio * (od[3 + k] - dg * c.SJ[k]);
                            c.path = 0;
                        }
                    }
                }
                __syncthreads();
            } else {
                const SplitWarp sw = c.sw;
                const float cf = c.cf;
                const bool all = c.sel0_all, wholestats = (MODE == 6 && a.plain);
                // residual form of the sums (as icgn_lean.hip): S1, S2 without
                // masked pixels (rf), T on the set-up's f side (rt)
                const bool rf = c.rform && !(MODE == 2 && c.fsel);
                const bool rt = !(MODE == 2 && c.fsel);
                const float kf = (float)c.kap;
                const float ck = (float)((1.0 - (double)kf) * (double)cf);
                const float c1k = (float)((1.0 + (double)kf) * (double)cf);
                float acc[6 + NJ];
#pragma unroll
                for (int k = 0; k < 6 + NJ; ++k) acc[k] = 0.f;
                for (int p = tid; p < npix; p += BS) {
                    int i, j;
                    pix_ij(p, nx, rcp, i, j);
                    const float dx = (float)i - hx, dy = (float)j - hy;
                    float g = 0.f;
                    if (use_tile) {
                        g = split_gather(sw, tile, tnx, dx, dy);
                    } else {
                        const float su = fmaf(sw.p1, dx, fmaf(sw.p3, dy, sw.fu));
                        const float sv = fmaf(sw.p2, dx, fmaf(sw.p4, dy, sw.fv));
                        const float fsu = floorf(su), fsv = floorf(sv);
                        float wu[4], wv[4];
                        keys_w2(__builtin_amdgcn_fractf(su), wu);
                        keys_w2(__builtin_amdgcn_fractf(sv), wv);
                        const int iu = (int)((dx + sw.cu) + fsu), iv = (int)((dy + sw.cv) + fsv);
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            const int y = min(max(tv0 + iv + q, 1), fr.W);
                            const float *col = fr.g + (size_t)(y - 1) * fr.H;
                            float s = 0.f;
#pragma unroll
                            for (int r = 0; r < 4; ++r) s = fmaf(wu[r], gld(col, (size_t)(min(max(tu0 + iu + r, 1), fr.H) - 1)), s);
                            g = fmaf(wv[q], s, g);
                        }
                        g *= 0.25f;  // the weights are 2 w
                    }
                    gbuf[p] = g;
                    const bool s = all || sel0[p];
                    const int x = xa + i, y = ya + j;
                    // f on the set-up's f side, as the set-up summed it
                    float fvv = 0.f;
                    if (s) {
                        fvv = fidx(fr.f, fr.H, x, y);
                        if (!(MODE == 6 && a.plain)) fvv *= fidx(fr.mask, fr.H, x, y);
                    }
                    const float d = g - cf;
                    const float r = fmaf(-kf, fvv, g) - ck;
                    const float qv = fmaf(kf, fvv, g) - c1k;
                    const float s1 = rf ? r : d, s2 = rf ? qv : d, t = rt ? r : d;
                    const bool nz = wholestats || fabsf(g) > 1e-10f;
                    acc[0] += nz ? 1.f : 0.f;
                    acc[1] += nz ? s1 : 0.f;
                    acc[2] = nz ? fmaf(s1, s2, acc[2]) : acc[2];
                    acc[3] += g == 0.f ? 1.f : 0.f;
                    acc[4] += isnan(g) ? 1.f : 0.f;
                    const size_t gi = (size_t)((x - 4) + (y - 4) * Hd);
                    const float fxv = s ? gld(fr.dfdx, gi) : 0.f, fyv = s ? gld(fr.dfdy, gi) : 0.f;
                    const float u = fxv * t, v = fyv * t;
                    if (MODE == 6) {
                        acc[6] = fmaf(dx, u, acc[6]);
                        acc[7] = fmaf(dx, v, acc[7]);
                        acc[8] = fmaf(dy, u, acc[8]);
                        acc[9] = fmaf(dy, v, acc[9]);
                        acc[10] += u;
                        acc[11] += v;
                    } else {
                        acc[6] += u;
                        acc[7] += v;
                    }
                }
                float o[6 + NJ];
                block_sum<BS, 6 + NJ>(acc, redf, o);
                if (tid == 0) {
                    c.path = 0;
                    const bool quad = !(MODE == 6 && a.plain);
                    if (quad && o[4] > 0) {
                        c.status = ST_ERROR;  // logical(NaN) throws
                        c.state = 2;
                    } else if (quad && (c.count > 0 || o[3] > 0)) {
                        c.path = 1;
                    } else if (MODE == 2 && c.fsel) {
                        c.path = 2;
                    } else {
                        // gbar - c = S1/n, bottomg^2 = S2 - S1^2/n; b = FJ - (bf/bg) (T - (gbar-c) SJ)
                        const double n = o[0];
                        double dg = 0, var = 0, T[NJ];
                        bool form_ok = true;
                        if (rf) {
                            // the residual form takes the nonzero set of g as the whole window
                            float sums[2 + NJ];
                            sums[0] = o[1];
                            sums[1] = o[2];
                            for (int k = 0; k < NJ; ++k) sums[2 + k] = o[6 + k];
                            if (n == (double)npix)
                                residual_stats<NJ>(n, sums, kf, ck, c1k, cf, c.F1c, c.F2c, c.FJc, c.SJ,
                                                   dg, var, T);
                            else
                                form_ok = false;
                        } else {
                            dg = o[1] / n;
                            var = o[2] - o[1] * dg;
                            const double k = kf;
                            const double corr = (double)ck - (1.0 - k) * (double)cf;
                            for (int q = 0; q < NJ; ++q)
                                T[q] = rt ? (double)o[6 + q] + k * c.FJc[q] + corr * c.SJ[q] : (double)o[6 + q];
                        }
                        const double bf = c.bottomf;
                        if (!form_ok || !(n >= 2) || !(bf > 0) || !(var > 0) || !isfinite(bf) ||
                            !isfinite(var) || !isfinite(dg)) {
                            c.path = 2;  // literal per-pixel residual for degenerate statistics
                        } else {
                            const double sv = sqrt(var);
                            const double ratio = bf / sv;
                            c.kap = sv / bf;
                            for (int k = 0; k < NJ; ++k) c.bvec[k] = c.FJ[k] - ratio * (T[k] - dg * c.SJ[k]);
                        }
                    }
                }
                __syncthreads();
            }

            if (c.state == 0 && c.path != 0) {
                // ------------------------------------- exact (masked) path
                const bool masked = c.path == 1;
                const bool fsel = MODE == 2 && (masked || c.fsel);
                if (masked) {
                    for (int p = tid; p < npix; p += BS) bwb[p] = gbuf[p] != 0.f;
                    __syncthreads();
                    // 6-DOF: 8-connected (:171); 2-DOF: 4-connected (:160)
                    flood_fill<BS>(bwb, sel, nx, ny, seed_r, seed_c, MODE == 6 ? 8 : 4, npix, rcp);
                    int allsel = 1;
                    for (int p = tid; p < npix; p += BS) allsel &= sel[p];
                    allsel = __syncthreads_and(allsel);
                    if (allsel) {
                        // the region is the whole window: use the window mask
                        // instead, 8-connected (:175-176, Subpb1 :164-165)
                        for (int p = tid; p < npix; p += BS) {
                            int i, j;
                            pix_ij(p, nx, rcp, i, j);
                            tm[p] = fidx(fr.mask, fr.H, xa + i, ya + j) != 0.f;
                        }
                        __syncthreads();
                        flood_fill<BS>(tm, sel, nx, ny, seed_r, seed_c, 8, npix, rcp);
                    }
                    for (int p = tid; p < npix; p += BS)
                        if (!sel[p]) gbuf[p] = 0.f;
                    __syncthreads();
                }
                // one pass: every FP64 sum of the iteration about c = cf
                // (g statistics over the nonzero g of the region, T = sum J (g - c)
                // over the f side; the 2-DOF region's f statistics, FJ, SJ and H
                // when f is rebuilt on it); the literal passes below take over
                // when the statistics are degenerate (NaN and infinity as the
                // reference propagates them)
                {
                    constexpr int K = 4 + NJ + 10;
                    double v[K];
#pragma unroll
                    for (int k = 0; k < K; ++k) v[k] = 0.0;
                    const bool whole = MODE == 6 && a.plain;
                    const bool all = c.sel0_all;
                    const double cd = c.cf;
                    for (int p = tid; p < npix; p += BS) {
                        int i, j;
                        pix_ij(p, nx, rcp, i, j);
                        const int x = xa + i, y = ya + j;
                        const double g = gbuf[p];
                        const bool nz = whole || fabs(g) > 1e-10;
                        const double e = g - cd;
                        v[0] += nz ? 1.0 : 0.0;
                        v[1] += nz ? e : 0.0;
                        v[2] += nz ? e * e : 0.0;
                        v[3] += masked ? (double)sel[p] : 0.0;
                        const bool s = fsel ? (bool)sel[p] : (all || sel0[p]);
                        if (s) {
                            const size_t gi = (size_t)((x - 4) + (y - 4) * Hd);
                            const float2 gq = grad_px(gl, gradf, fr, p, gi);
                            const double fx = gq.x, fy = gq.y;
                            if (MODE == 6) {
                                const double dx = i - (double)hx, dy = j - (double)hy;
                                v[4] += dx * fx * e;
                                v[5] += dx * fy * e;
                                v[6] += dy * fx * e;
                                v[7] += dy * fy * e;
                                v[8] += fx * e;
                                v[9] += fy * e;
                            } else {
                                v[4] += fx * e;
                                v[5] += fy * e;
                                if (fsel) {
                                    const double fv = (double)fidx(fr.f, fr.H, x, y) *
                                                      (double)fidx(fr.mask, fr.H, x, y);
                                    const bool fnz = fabs(fv) > 1e-10;
                                    const double ef = fv - cd;
                                    v[6] += fnz ? 1.0 : 0.0;
                                    v[7] += fnz ? ef : 0.0;
                                    v[8] += fnz ? ef * ef : 0.0;
                                    v[9] += fabs(fv);
                                    v[10] += fx;
                                    v[11] += fy;
                                    v[12] += fx * fv;
                                    v[13] += fy * fv;
                                    v[14] += fx * fx;
                                    v[15] += fx * fy;
                                    v[16] += fy * fy;
                                }
                            }
                        }
                    }
                    double o[K];
                    block_sum<BS, K>(v, redd, o);
                    if (tid == 0) {
                        const double ng = o[0], dg = o[1] / ng, varg = o[2] - o[1] * dg;
                        bool ok = ng >= 2 && varg > 0 && isfinite(varg) && isfinite(dg);
                        double mf = c.meanf, bf = c.bottomf;
                        double FJ[6], SJ[6];
                        for (int k = 0; k < NJ; ++k) {
                            FJ[k] = c.FJ[k];
                            SJ[k] = c.SJ[k];
                        }
                        if (MODE == 2 && fsel) {
                            const double nf = o[6], df = o[7] / nf, varf = o[8] - o[7] * df;
                            ok = ok && nf >= 2 && varf > 0 && isfinite(varf) && isfinite(df);
                            mf = cd + df;
                            bf = sqrt(varf);
                            SJ[0] = o[10];
                            SJ[1] = o[11];
                            FJ[0] = o[12] - mf * o[10];
                            FJ[1] = o[13] - mf * o[11];
                        }
                        ok = ok && bf > 0 && isfinite(bf);
                        c.lit = !ok;
                        if (ok) {
                            // pass A and A2's side effects
                            c.ng = ng;
                            c.gmean = cd + dg;
                            c.bottomg = sqrt(varg);
                            if (masked) c.sum_bw2 = (int)o[3];
                            if (masked && MODE == 2) c.fsel = 1;
                            if (MODE == 2 && fsel) {
                                c.nf = o[6];
                                c.meanf = mf;
                                c.bottomf = bf;
                                c.sabs = o[9];
                                c.Hm[0] = o[14];
                                c.Hm[1] = o[15];
                                c.Hm[2] = o[16];
                                const double s2 = 2.0 / (bf * bf);
                                const double H2a = o[14] * s2 + a.mu, H2b = o[15] * s2,
                                             H2d = o[16] * s2 + a.mu;
                                const double dl = 1e-3 * (H2a > H2d ? H2a : H2d);
                                c.tH[0] = H2a + dl;
                                c.tH[1] = c.tH[2] = H2b;
                                c.tH[3] = H2d + dl;
                            }
                            // b = bottomf sum J (fhat - ghat) = FJ - (bf/bg) (T - (gbar - c) SJ)
                            const double ratio = bf / c.bottomg;
                            for (int k = 0; k < NJ; ++k)
                                c.bvec[k] = FJ[k] - ratio * (o[4 + k] - dg * SJ[k]);
                            const double kp = c.bottomg / bf;
                            if (isfinite(kp) && kp > 0) c.kap = kp;
                        }
                    }
                    __syncthreads();
                }
                if (c.lit) {
                // pass A: counts and sums (g; 2-DOF masked also f and H)
                {
                    double v[9];
#pragma unroll
                    for (int k = 0; k < 9; ++k) v[k] = 0.0;
                    const bool whole = MODE == 6 && a.plain;
                    for (int p = tid; p < npix; p += BS) {
                        const double g = gbuf[p];
                        const bool nz = whole || fabs(g) > 1e-10;
                        v[0] += nz ? 1.0 : 0.0;
                        v[1] += nz ? g : 0.0;
                        v[2] += masked ? (double)sel[p] : 0.0;
                        if (fsel) {
                            int i, j;
                            pix_ij(p, nx, rcp, i, j);
                            const int x = xa + i, y = ya + j;
                            const bool s = sel[p];
                            const size_t gi = (size_t)((x - 4) + (y - 4) * Hd);
                            const float2 gq = s ? grad_px(gl, gradf, fr, p, gi) : make_float2(0.f, 0.f);
                            const double fx = gq.x, fy = gq.y;
                            const double fv = s ? (double)fidx(fr.f, fr.H, x, y) * (double)fidx(fr.mask, fr.H, x, y) : 0.0;
                            const bool fnz = fabs(fv) > 1e-10;
                            v[3] += fnz ? 1.0 : 0.0;
                            v[4] += fnz ? fv : 0.0;
                            v[5] += fx * fx;
                            v[6] += fx * fy;
                            v[7] += fy * fy;
                            v[8] += fabs(fv);
                        }
                    }
                    double o[9];
                    block_sum<BS, 9>(v, redd, o);
                    if (tid == 0) {
                        c.gmean = o[0] > 0 ? o[1] / o[0] : NAN;
                        c.ng = o[0];
                        c.nf = o[3];
                        if (masked) c.sum_bw2 = (int)o[2];
                        if (fsel) {
                            c.meanf = o[3] > 0 ? o[4] / o[3] : NAN;
                            c.sabs = o[8];
                            c.Hm[0] = o[5];
                            c.Hm[1] = o[6];
                            c.Hm[2] = o[7];
                        }
                    }
                    __syncthreads();
                }
                // pass A2: second moments about the means
                {
                    double v[2] = {0.0, 0.0};
                    const bool whole = MODE == 6 && a.plain;
                    const double mg = c.gmean, mf = c.meanf;
                    for (int p = tid; p < npix; p += BS) {
                        const double g = gbuf[p];
                        if (whole || fabs(g) > 1e-10) v[0] += (g - mg) * (g - mg);
                        if (fsel && sel[p]) {
                            int i, j;
                            pix_ij(p, nx, rcp, i, j);
                            const double fv = (double)fidx(fr.f, fr.H, xa + i, ya + j) *
                                              (double)fidx(fr.mask, fr.H, xa + i, ya + j);
                            if (fabs(fv) > 1e-10) v[1] += (fv - mf) * (fv - mf);
                        }
                    }
                    double o[2];
                    block_sum<BS, 2>(v, redd, o);
                    if (tid == 0) {
                        const double ng = c.ng;
                        c.bottomg = ng >= 2 ? sqrt(o[0]) : (ng == 1 ? 0.0 : NAN);
                        if (masked && MODE == 2) c.fsel = 1;
                        if (fsel) {
                            const double nf = c.nf;
                            const double bf = nf >= 2 ? sqrt(o[1]) : (nf == 1 ? 0.0 : NAN);
                            c.bottomf = bf;
                            const double s2 = 2.0 / (bf * bf);
                            const double H2a = c.Hm[0] * s2 + a.mu, H2b = c.Hm[1] * s2, H2d = c.Hm[2] * s2 + a.mu;
                            const double dl = 1e-3 * (H2a > H2d ? H2a : H2d);
                            c.tH[0] = H2a + dl;
                            c.tH[1] = c.tH[2] = H2b;
                            c.tH[3] = H2d + dl;
                        }
                    }
                    __syncthreads();
                }
                // pass B: b = bottomf * sum(J r), r evaluated on every pixel (:362-371)
                {
                    double v[NJ];
#pragma unroll
                    for (int k = 0; k < NJ; ++k) v[k] = 0.0;
                    const double mf = c.meanf, bf = c.bottomf, mg = c.gmean, bg = c.bottomg;
                    const bool all = c.sel0_all;
                    for (int p = tid; p < npix; p += BS) {
                        int i, j;
                        pix_ij(p, nx, rcp, i, j);
                        const int x = xa + i, y = ya + j;
                        bool s;
                        double fv;
                        if (fsel) {
                            s = sel[p];
                            fv = s ? (double)fidx(fr.f, fr.H, x, y) * (double)fidx(fr.mask, fr.H, x, y) : 0.0;
                        } else {
                            s = all || sel0[p];
                            fv = fidx(fr.f, fr.H, x, y);
                            if (!(MODE == 6 && a.plain)) fv *= (double)fidx(fr.mask, fr.H, x, y);
                            if (!s) fv = 0.0;
                        }
                        const size_t gi = (size_t)((x - 4) + (y - 4) * Hd);
                        const float2 gq = s ? grad_px(gl, gradf, fr, p, gi) : make_float2(0.f, 0.f);
                        const double fx = gq.x, fy = gq.y;
                        const double r = (fv - mf) / bf - ((double)gbuf[p] - mg) / bg;
                        if (MODE == 6) {
                            const double dx = i - (double)hx, dy = j - (double)hy;
                            v[0] += dx * fx * r;
                            v[1] += dx * fy * r;
                            v[2] += dy * fx * r;
                            v[3] += dy * fy * r;
                            v[4] += fx * r;
                            v[5] += fy * r;
                        } else {
                            v[0] += fx * r;
                            v[1] += fy * r;
                        }
                    }
                    double o[NJ];
                    block_sum<BS, NJ>(v, redd, o);
                    if (tid == 0) {
                        for (int k = 0; k < NJ; ++k) c.bvec[k] = c.bottomf * o[k];
                        const double kp = c.bottomg / c.bottomf;
                        if (isfinite(kp) && kp > 0) c.kap = kp;
                    }
                    __syncthreads();
                }
                }  // literal passes
            }

            // ------------------------------- norm test, GN solve, IC update
            if (tid == 0 && c.state == 0) {
                const double tol = a.tol;
                double b[6];
                for (int k = 0; k < NJ; ++k) b[k] = c.bvec[k];
                double d[6] = {0, 0, 0, 0, 0, 0};
                double nb2 = 0.0;
                if (MODE == 6) {
                    for (int k = 0; k < 6; ++k) nb2 += b[k] * b[k];
                } else {
                    // :379 tempb = 2 b / bottomf^2 + mu (P - UOld - vdual)
                    const double bf = c.bottomf, s2 = 2.0 / (bf * bf);
                    const long idx = c.idx;
                    b[0] = b[0] * s2 + a.mu * (c.P[4] - a.UOld[2 * idx] - a.vdual[2 * idx]);
                    b[1] = b[1] * s2 + a.mu * (c.P[5] - a.UOld[2 * idx + 1] - a.vdual[2 * idx + 1]);
                    nb2 = b[0] * b[0] + b[1] * b[1];
                }
                double nn = sqrt(nb2);
                c.normAbs = nn;
                if (c.it == 1) c.normInit = nn;
                nn = c.normInit > tol ? nn / c.normInit : 0.0;
                c.normNew = nn;
                const double tol2 = tol * tol;
                if (near_tol(nb2, tol2, a.guard2_abs) ||
                    (c.normInit > tol && near_tol(nn * nn, tol2, a.guard2))) {
                    c.exact = 1;  // the test is within FP32 reach of tol: FP64 re-solve
                    c.state = 3;
                } else if (nn < tol || c.normAbs < tol) {
                    c.state = 1;
                } else {
                    if (MODE == 6) {
                        const double nbv[6] = {-b[0], -b[1], -b[2], -b[3], -b[4], -b[5]};
                        solve_factored<6>(c.Hm, c.piv, c.fkind, nbv, d);
                    } else {
                        const double nbv[2] = {-b[0], -b[1]};
                        double d2[2];
                        solve_small<2>(c.tH, nbv, d2);
                        d[4] = d2[0];
                        d[5] = d2[1];
                    }
                    if (!ic_compose(c.P, d)) c.state = 1;
                }
            }
            __syncthreads();
        }

        // ---------------------------------------------------------- outputs
        if (tid == 0 && c.state == 3) {
            a.status[c.idx] = c.exact ? ST_EXACT : ST_DEFER;
            defer_append(a, c.idx, a.wx[c.idx], MODE == 2 ? a.wy[c.idx] : a.wx[c.idx]);
            if (c.exact && a.exact_count) atomicAdd(a.exact_count, 1);
        }
        if (tid == 0 && c.state <= 2) {
            int st = c.status;
            if (c.state == 1) {
                const double tol = a.tol;
                int it = c.it;
                if (!(c.normNew < tol || c.normAbs < tol)) it = a.max_it + 1;
                if (isnan(c.normNew)) it = a.max_it + 1;
                if (!(MODE == 6 && a.plain) && c.sabs < 1e-6) it = a.max_it + 3;
                st = it;
            }
            const long idx = c.idx;
            double P[6];
            for (int k = 0; k < 6; ++k) P[k] = c.P[k];
            if (st == ST_ERROR)
                for (int k = 0; k < 6; ++k) P[k] = NAN;
            if (st == ST_MASKED && MODE == 6) P[0] = P[1] = P[2] = P[3] = 0.0;
            if (st == ST_MASKED && MODE == 6) {
                P[4] = a.U0[2 * idx];
                P[5] = a.U0[2 * idx + 1];
            }
            a.U[2 * idx] = P[4];
            a.U[2 * idx + 1] = P[5];
            if (MODE == 6) {
                for (int k = 0; k < 4; ++k) a.F[4 * idx + k] = P[k];
                const bool zeroH = st == ST_ERROR || st == ST_MASKED;
                for (int k = 0; k < 21; ++k) a.H21[idx + k * a.N] = zeroH ? 0.0 : c.H21[k];
            }
            a.status[idx] = st;
        }
        __syncthreads();
        if (!a.work_counter) break;
    }
}

// Host-side launch helpers (instantiated for the block sizes the autotuner may pick).
size_t icgn_lds_bytes(int npix_cap, int tile_cap, bool exact, bool grad)
{
    // g per pixel (FP64 in the re-solve), DfDx and DfDy per pixel when staged,
    // the g tile, four byte masks per pixel
    size_t b = (exact ? sizeof(double) : sizeof(float)) * (size_t)npix_cap +
               (exact && grad ? sizeof(float2) * (size_t)npix_cap : 0) +
               sizeof(float) * (size_t)tile_cap + 4u * (size_t)npix_cap;
    return (b + 15) & ~(size_t)15;
}

template <int BS, int MODE, bool EXACT>
static hipError_t launch_bs(const IcgnArgs &a, int grid, hipStream_t s)
{
    const size_t lds = icgn_lds_bytes(a.npix_cap, a.tile_cap, EXACT, a.grad_lds != 0);
    static bool attr_set = false;
    if (!attr_set) {
        hipFuncSetAttribute((const void *)icgn_kernel<BS, MODE, EXACT>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 8192);
        attr_set = true;
    }
    hipLaunchKernelGGL((icgn_kernel<BS, MODE, EXACT>), dim3(grid), dim3(BS), lds, s, a);
    return hipGetLastError();
}

hipError_t icgn_launch(const IcgnArgs &a, int mode, int bs, int grid, hipStream_t s)
{
    if (mode == 6) {
        switch (bs) {
        case 64: return launch_bs<64, 6, false>(a, grid, s);
        case 128: return launch_bs<128, 6, false>(a, grid, s);
        case 256: return launch_bs<256, 6, false>(a, grid, s);
        case 512: return launch_bs<512, 6, false>(a, grid, s);
        }
    } else {
        switch (bs) {
        case 64: return launch_bs<64, 2, false>(a, grid, s);
        case 128: return launch_bs<128, 2, false>(a, grid, s);
        case 256: return launch_bs<256, 2, false>(a, grid, s);
        case 512: return launch_bs<512, 2, false>(a, grid, s);
        }
    }
    return hipErrorInvalidValue;
}

hipError_t icgn_exact_launch(const IcgnArgs &a, int mode, int bs, int grid, hipStream_t s)
{
    if (bs == 64) return mode == 6 ? launch_bs<64, 6, true>(a, grid, s) : launch_bs<64, 2, true>(a, grid, s);
    return mode == 6 ? launch_bs<256, 6, true>(a, grid, s) : launch_bs<256, 2, true>(a, grid, s);
}

}  // namespace staq
