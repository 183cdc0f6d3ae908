// Lean wavefront-per-subset IC-GN kernel for subsets whose window holds no
// masked pixel (gfx950).  It is the unmasked path of icgn_wave.hip and nothing
// else, which keeps its register file small (several waves per SIMD instead of
// one, so that the dependent chains of many subsets' iterations share each
// SIMD) and its iteration short (fewer vector instructions per pixel):
//
//   * no bwselect regions, no window growth, no FP64 residual pass: a subset
//     that meets any of them (a zero of f*mask in its window, a zero of the
//     warped g under the quadtree rules, degenerate statistics, a Hessian that
//     is not positive definite, a window beyond 64 * PPT * NW pixels, a warped
//     support beyond the g tile) is left for the fallback launch (status -2),
//     which solves it from the start;
//   * per pixel slot the lane keeps df/dx, df/dy and the window offsets
//     (x - x0, y - y0) in registers, so an iteration is the warp, the Keys
//     gather from the resident g tile, eight sums, and the FP64 step;
//   * the sums, the statistics, the Gauss-Newton step and the compose are
//     those of icgn_wave.hip (shared through wave_common.h) up to rounding:
//     this kernel forms r and q at four times their value (exact), takes the
//     warped window's extremes in closed form and multiplies by 1/n and
//     1/bottomf instead of dividing, so both kernels hold the same bar
//     against the FP64 evaluation, not bit-identical results.
//
// Semantics: funICGNQuadtree.m:40-451 / funICGN.m:40-323 (6-DOF) and
// funICGN_Subpb1.m:40-474 (2-DOF) on windows without masked pixels.
#include "wave_common.h"

namespace staq {

namespace {

// Window offsets (dx, dy) of one pixel slot in one register: two FP16 values,
// exact for the integers and half-integers of windows up to 2^10 pixels wide.
__device__ __forceinline__ unsigned pack_dd(float dx, float dy)
{
    const unsigned lo = __builtin_bit_cast(unsigned short, (_Float16)dx);
    const unsigned hi = __builtin_bit_cast(unsigned short, (_Float16)dy);
    return lo | (hi << 16);
}

__device__ __forceinline__ void unpack_dd(unsigned v, float &dx, float &dy)
{
    dx = (float)__builtin_bit_cast(_Float16, (unsigned short)(v & 0xffffu));
    dy = (float)__builtin_bit_cast(_Float16, (unsigned short)(v >> 16));
}

template <int NW>
struct LeanLds {
    double Lp[21];   // packed Cholesky factor of H (column-major lower triangle)
    double Ld[6];    // 1 / diag(Lp)
    double H2i[4];   // 2-DOF: inverse of the damped 2x2 system
    float Hs[21];    // H as summed (FP32), for its lane-parallel store
    double FJ[6], SJ[6];
    // sums about cf = (float) meanf, for the residual form of the
    // per-iteration sums: FJc = sum J (f - cf), F1c = sum (f - cf),
    // F2c = sum (f - cf)^2
    double FJc[6], F1c, F2c;
    // exchange between the wavefronts of one subset (double-buffered)
    float xf[2][NW][40];
    double xd[2][NW][16];
    int xi[2][NW][8];
    // eight or more wavefronts: the Gauss-Newton step of wavefront 0
    double Pn[6], rel2, abs2, inv_init2, kap;
    int stop, init_ok;
};

template <int PPT, int MODE, int NW>
// two wavefronts per subset: three waves per SIMD (the register allocator
// otherwise settles on two, which measured slower on 33 x 33 windows); other
// forms: the allocator's own choice (a three-wave cap measured slower there)
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(NW == 2 ? 3 : 1)))
void icgn_lean_kernel(IcgnArgs a)
{
    constexpr int NJ = MODE;
    constexpr int CAP = 64 * PPT * NW;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ LeanLds<NW> L;
    __shared__ long item_sh;
    float *tile = reinterpret_cast<float *>(smem);
    // f of every pixel slot (f * mask under the quadtree rules), read each
    // iteration by the residual form of the sums; after the g tile
    float *fwin = tile + a.tile_cap;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);  // scalar: alike in every lane
    const bool quad = !(MODE == 6 && a.plain);
    int par = 0;
    // pixel of slot k for lane index l.  The lane index reaches each use
    // through fresh_lane(), an empty asm the compiler cannot see through, so
    // that what derives from it (pixel offsets, LDS addresses) is recomputed
    // where it is used instead of being kept live, and spilled, across the
    // iteration loop; the asm emits no instruction.
    auto pixel_of = [&](int l, int k) { return l + 64 * (k * NW + wv); };
    auto fresh_lane = [&]() {
        int l = lane;
        asm volatile("" : "+v"(l));
        return l;
    };

    for (long pass = 0;; ++pass) {
        if (tid == 0)
            item_sh = a.work_counter ? a.item0 + atomicAdd(a.work_counter, 1)
                                     : (pass == 0 ? a.item0 + blockIdx.x : a.item0 + a.n_items);
        __syncthreads();
        const long item = item_sh;
        if (item >= a.item0 + a.n_items) break;
        const long idx = a.order ? gld(a.order, (size_t)(item)) : item;
        const int lane_s = fresh_lane();  // for this subset's set-up
        auto pixel = [&](int k) { return pixel_of(lane_s, k); };
        const int slot = a.slot ? gld(a.slot, (size_t)(idx)) : 0;
        const FrameDev fr = load_frame(a.frames, slot);
        const int Hd = fr.H - 6;

        double x0 = gld(a.coords, (size_t)(idx)), y0 = gld(a.coords, (size_t)(idx + a.N));
        if (!a.plain) {
            x0 = round(x0);  // LocalICGNQuadtree.m:69
            y0 = round(y0);
        }
        const int wx = gld(a.wx, (size_t)(idx));
        const int wy = MODE == 2 ? gld(a.wy, (size_t)(idx)) : wx;
        double P[6];
        if (MODE == 6) {
            P[0] = P[1] = P[2] = P[3] = 0.0;
            P[4] = gld(a.U0, (size_t)(2 * idx));
            P[5] = gld(a.U0, (size_t)(2 * idx + 1));
        } else {
            for (int k = 0; k < 4; ++k) P[k] = gld(a.FOld, (size_t)(4 * idx + k));
            P[4] = gld(a.UOld, (size_t)(2 * idx));
            P[5] = gld(a.UOld, (size_t)(2 * idx + 1));
        }
        int state = 0, status = 0;  // 0 run, 1 loop done, 2 final status, 3 deferred
        bool exact = false;  // deferred to the FP64 re-solve (guard band, ill-conditioned H)
        {
            const double xa = x0 - wx * 0.5, xb = x0 + wx * 0.5, ya = y0 - wy * 0.5,
                         yb = y0 + wy * 0.5;
            const bool fits = xa == floor(xa) && ya == floor(ya) && xb == floor(xb) &&
                              yb == floor(yb) && xa >= 4 && ya >= 4 && xb <= fr.H - 3 &&
                              yb <= fr.W - 3;
            if (!fits) {
                state = 2;
                status = ST_ERROR;
            } else if ((wx + 1) * (wy + 1) > CAP) {
                state = 3;
            }
        }

        // ------------------------------------------------------------ setup
        const int nx = wx + 1, ny = wy + 1, npix = nx * ny;
        const int xa = (int)x0 - wx / 2, ya = (int)y0 - wy / 2;
        const float rnx = 1.0f / (float)nx;
        const float hx = wx * 0.5f, hy = wy * 0.5f;
        // per-slot df/dx, df/dy and the window offsets (dx, dy) as a pair of
        // FP16 values (exact: integers or half-integers below 2^10); f of
        // every slot sits in LDS.  f statistics and steepest-descent sums in
        // one pass about c0 = f at the centre
        float fxr[PPT], fyr[PPT];
        unsigned ddr[PPT];
        unsigned vbits = 0u;  // bit k: slot k holds a window pixel
        double meanf = 0, bottomf = 0, sabs = 0, rel2 = 1, abs2 = 1, inv_init2 = 0;
        bool init_ok = false;
        const double tol2 = a.tol * a.tol;
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            // a slot past the window repeats the last pixel and adds nothing
            const int p = pixel(k);
            const int pc = p < npix ? p : npix - 1;
            const int j = (int)(((float)pc + 0.5f) * rnx), i = pc - j * nx;
            ddr[k] = pack_dd((float)i - hx, (float)j - hy);
            fxr[k] = fyr[k] = 0.f;
            vbits |= p < npix ? (1u << k) : 0u;
            if (p >= npix) fwin[p] = 0.f;
        }
        if (state == 0) {
            const float c0 = fidx(fr.f, fr.H, xa + nx / 2, ya + ny / 2) *
                             (quad ? fidx(fr.mask, fr.H, xa + nx / 2, ya + ny / 2) : 1.f);
            // the masked-pixel count (funICGNQuadtree.m:61 / funICGN.m:54) in
            // the same pass; a window with masked pixels leaves the sums unused
            int cnt = 0;
            float s_n = 0.f, s_1 = 0.f, s_2 = 0.f, s_abs = 0.f;
            float sj[NJ], sjf[NJ];
            constexpr int NH = MODE == 6 ? 21 : 3;
            float hs[NH];
#pragma unroll
            for (int q = 0; q < NJ; ++q) sj[q] = sjf[q] = 0.f;
#pragma unroll
            for (int q = 0; q < NH; ++q) hs[q] = 0.f;
#pragma unroll
            for (int k = 0; k < PPT; ++k) {
                const int p = pixel(k);
                if (p < npix) {
                    const int j = (int)(((float)p + 0.5f) * rnx), i = p - j * nx;
                    const int x = xa + i, y = ya + j;
                    const size_t gi = (size_t)((x - 4) + (y - 4) * Hd);
                    fxr[k] = gld(fr.dfdx, gi);
                    fyr[k] = gld(fr.dfdy, gi);
                    float f = fidx(fr.f, fr.H, x, y);
                    if (quad) f *= fidx(fr.mask, fr.H, x, y);
                    fwin[p] = f;
                    const float fx = fxr[k], fy = fyr[k];
                    cnt += (MODE == 6 && a.plain) ? fx == 0.f : f == 0.f;
                    const bool nz = !quad || fabsf(f) > 1e-10f;
                    const float df = f - c0;
                    s_n += nz ? 1.f : 0.f;
                    s_1 += nz ? df : 0.f;
                    s_2 += nz ? df * df : 0.f;
                    s_abs += fabsf(f);
                    if (MODE == 6) {
                        const float dx = (float)i - hx, dy = (float)j - hy;
                        const float J[6] = {fx * dx, fy * dx, fx * dy, fy * dy, fx, fy};
                        int t = 0;
#pragma unroll
                        for (int c = 0; c < 6; ++c) {
                            sj[c] += J[c];
                            sjf[c] += J[c] * df;
#pragma unroll
                            for (int r = c; r < 6; ++r) hs[t++] += J[r] * J[c];
                        }
                    } else {
                        sj[0] += fx;
                        sj[1] += fy;
                        sjf[0] += fx * df;
                        sjf[1] += fy * df;
                        hs[0] += fx * fx;
                        hs[1] += fx * fy;
                        hs[2] += fy * fy;
                    }
                }
            }
            // every sum over the window in one exchange (FP32 totals)
            constexpr int NS = NH + 4 + 2 * NJ;
            float sf[NS];
#pragma unroll
            for (int q = 0; q < NH; ++q) sf[q] = hs[q];
            sf[NH] = s_n;
            sf[NH + 1] = s_1;
            sf[NH + 2] = s_2;
            sf[NH + 3] = s_abs;
#pragma unroll
            for (int q = 0; q < NJ; ++q) {
                sf[NH + 4 + q] = sj[q];
                sf[NH + 4 + NJ + q] = sjf[q];
            }
            int ci[1] = {(int)wave_sum_dpp((float)cnt)};
            group_reduce<NW, NS, 0, 1>(L, par, sf, nullptr, ci);
#pragma unroll
            for (int q = 0; q < NH; ++q) hs[q] = sf[q];
            double sd[4 + 2 * NJ];
#pragma unroll
            for (int q = 0; q < 4 + 2 * NJ; ++q) sd[q] = sf[NH + q];
            {
                const int count = ci[0];
                const double n0 = (double)nx * ny;
                const double cap = (MODE == 6 && a.plain) ? 0.5 * n0 : 0.4 * n0;
                if (!(count < cap)) {
                    state = 2;
                    status = ST_MASKED;
                } else if ((MODE == 6 && a.plain) ? count > 0.1 * n0 : count > 0) {
                    state = 3;  // window growth and mask regions: the general kernel
                }
            }
            const double n = sd[0];
            const double S1 = sd[1], S2 = sd[2];
            sabs = sd[3];
            meanf = n > 0 ? (double)c0 + S1 / n : NAN;
            bottomf = n >= 2 ? sqrt(fmax(S2 - S1 * S1 / n, 0.0)) : (n == 1 ? 0.0 : NAN);
            // the residual form takes every window pixel as nonzero f: a
            // window with |f| <= 1e-10 somewhere goes to the general kernel
            if (n != (double)npix) state = 3;
            const double dm = (double)c0 - meanf;
            // the same sums about cf = (float) meanf (exact shifts in FP64)
            const double e0 = (double)c0 - (double)(isfinite(meanf) ? (float)meanf : 0.f);
            if (tid == 0) {
#pragma unroll
                for (int q = 0; q < NJ; ++q) {
                    L.SJ[q] = sd[4 + q];
                    L.FJ[q] = sd[4 + NJ + q] + dm * sd[4 + q];  // sum J (f - meanf)
                    L.FJc[q] = sd[4 + NJ + q] + e0 * sd[4 + q];  // sum J (f - cf)
                }
                L.F1c = S1 + n * e0;
                L.F2c = S2 + 2.0 * e0 * S1 + n * e0 * e0;
            }
            if (MODE == 6) {
                double h[21];
#pragma unroll
                for (int q = 0; q < 21; ++q) h[q] = hs[q];
                if (tid == 0)
                    for (int q = 0; q < 21; ++q) L.Hs[q] = hs[q];
                if (chol6(h)) {
                    if (tid == 0) {
                        for (int q = 0; q < 21; ++q) L.Lp[q] = h[q];
                        for (int c = 0; c < 6; ++c) L.Ld[c] = 1.0 / h[c * 6 - (c * (c - 1)) / 2];
                    }
                    double dg6[6];
#pragma unroll
                    for (int c = 0; c < 6; ++c) dg6[c] = h[c * 6 - (c * (c - 1)) / 2];
                    if (a.guard2 > 0 && pivots_spread(dg6)) {
                        state = 3;  // ill-conditioned H: the FP64 re-solve
                        exact = true;
                    }
                } else {
                    state = 3;  // LU of an indefinite H: the general kernel (or the FP64 re-solve)
                    exact = a.guard2 > 0;
                }
            } else {
                set_h2(L, hs[0], hs[1], hs[2], bottomf, a.mu);
            }
            __syncthreads();
            // H to HBM, one entry per lane (funICGNQuadtree.m:448)
            if (MODE == 6 && tid < 21) a.H21[idx + (long)tid * a.N] = (double)L.Hs[tid];
        }

        // ---------------------------------------------------------- iterate
        // 1/n and 1/bottomf, FP64 divisions once per subset
        const double inv_n = 1.0 / (double)npix, inv_bf = 1.0 / bottomf;
        int it = 0;
        int tu0 = 0, tv0 = 0, tnx = 0, tny = 0;  // resident g tile (1-based origin, extents)
        bool have_tile = false;
        double kap = 1.0;  // bottomg / bottomf of the last iteration
        const float cf = isfinite(meanf) ? (float)meanf : 0.f;
        double uo_u = 0.0, uo_v = 0.0, vd_u = 0.0, vd_v = 0.0;  // 2-DOF penalty targets
        if (MODE == 2) {
            uo_u = gld(a.UOld, (size_t)(2 * idx));
            uo_v = gld(a.UOld, (size_t)(2 * idx + 1));
            vd_u = gld(a.vdual, (size_t)(2 * idx));
            vd_v = gld(a.vdual, (size_t)(2 * idx + 1));
        }
        // state is the same in every lane; saying so keeps the iteration a
        // uniform loop (scalar branches), which the compiler cannot prove
        // through the lane-0 writes of the set-up
        state = __builtin_amdgcn_readfirstlane(state);
        while (state == 0) {
            if (!(it <= a.max_it && rel2 > tol2 && abs2 > tol2)) {
                state = 1;
                break;
            }
            ++it;
            bool nanP = false;
            for (int k = 0; k < 6; ++k) nanP |= isnan(P[k]);
            if (nanP) {
                if (MODE == 2 || quad) {
                    state = 2;
                    status = ST_ERROR;
                } else {
                    rel2 = it == 1 ? 0.0 : NAN;
                    abs2 = NAN;
                    state = 1;
                }
                break;
            }
            // bounds of the warped window from its corners (funICGNQuadtree.m:149-157)
            double umin, umax, vmin, vmax;
            bool out;
            {
                // the corners (x0 +- hx, y0 +- hy) warp to c +- A +- B, so the
                // extremes of each coordinate are c +- (|A| + |B|); alike in
                // every lane (no cross-lane step on the critical path)
                const double cu = x0 + P[4], cv = y0 + P[5];
                const double eu = fabs((1 + P[0]) * hx) + fabs(P[2] * hy);
                const double ev = fabs(P[1] * hx) + fabs((1 + P[3]) * hy);
                umin = cu - eu;
                umax = cu + eu;
                vmin = cv - ev;
                vmax = cv + ev;
                out = (umin < 3) | (umax > fr.H - 2) | (vmin < 3) | (vmax > fr.W - 2);
                // the test can only come out otherwise under the FP32 sums'
                // effect on P where an extreme lies near its own bound
                const bool edge = (fabs(umin - 3.0) < 1e-3) | (fabs(umax - (fr.H - 2.0)) < 1e-3) |
                                  (fabs(vmin - 3.0) < 1e-3) | (fabs(vmax - (fr.W - 2.0)) < 1e-3);
                if (edge && a.guard2 > 0) {
                    state = 3;  // the bounds test decided within FP32 reach: FP64 re-solve
                    exact = true;
                    break;
                }
            }
            if (out) {
                rel2 = 1e12;  // normOfWNew = 1e6 (:156)
                state = 1;
                break;
            }
            // support of this iteration's taps, one pixel of slack for the
            // FP32 local coordinates; the tile keeps a margin of up to 4
            const int nu0 = (int)floor(umin) - 2, nv0 = (int)floor(vmin) - 2;
            const int nu1 = (int)floor(umax) + 3, nv1 = (int)floor(vmax) + 3;
            const bool reload = !have_tile || nu0 < tu0 || nv0 < tv0 || nu1 >= tu0 + tnx ||
                                nv1 >= tv0 + tny;
            if (reload) {
                int m = 4;
                while (m > 0 && (nu1 - nu0 + 1 + 2 * m) * (nv1 - nv0 + 1 + 2 * m) > a.tile_cap) --m;
                tu0 = nu0 - m;
                tv0 = nv0 - m;
                tnx = nu1 - nu0 + 1 + 2 * m;
                tny = nv1 - nv0 + 1 + 2 * m;
                if (tnx * tny > a.tile_cap) {
                    // a warped support beyond the tile is left to the general
                    // kernel, which gathers from HBM; without a second gather
                    // path this kernel needs a third fewer registers
                    state = 3;
                    break;
                }
                have_tile = true;
            }
            // warp coordinates split into integer and FP32 fractional parts
            const SplitWarp sw = split_warp(P, x0, y0, tu0, tv0);
            if (reload) {
                __syncthreads();  // the last iteration's gathers are done with the old tile
                const int total = tnx * tny;
                const float rcpn = 1.0f / tnx;
                for (int t0 = 0; t0 < total; t0 += 64 * NW * 8) {
                    float v[8];
                    int o[8];
#pragma unroll
                    for (int q = 0; q < 8; ++q) {
                        const int t = t0 + tid + 64 * NW * q;
                        o[q] = t;
                        v[q] = 0.f;
                        if (t < total) {
                            const int sj = (int)(((float)t + 0.5f) * rcpn), ri = t - sj * tnx;
                            const int y = min(max(tv0 + sj, 1), fr.W), x = min(max(tu0 + ri, 1), fr.H);
                            v[q] = gld(fr.g, (size_t)((x - 1) + (y - 1) * fr.H));
                        }
                    }
#pragma unroll
                    for (int q = 0; q < 8; ++q)
                        if (o[q] < total) tile[o[q]] = v[q];
                }
                __syncthreads();
            }
            // Residual form of the sums, about kf * f: with r = g - kf f - ck
            // and q = g + kf f - c1k (ck = (1 - kf) cf, c1k = (1 + kf) cf
            // rounded to FP32), sum r, sum r q and sum J_q r are sums of small
            // terms near convergence (kf = bottomg / bottomf of the last
            // iteration), so their FP32 rounding stays far below the
            // tolerance the stopping test resolves; the sums about cf follow
            // exactly in FP64 from the set-up's sums of f (funICGNQuadtree.m:
            // 269-270, 362-371).  Every lane runs every slot (no branch splits
            // the unrolled slots, so their gathers overlap).
            const float kf = (float)kap;
            const float ck = (float)((1.0 - (double)kf) * (double)cf);
            const float c1k = (float)((1.0 + (double)kf) * (double)cf);
            // the gather returns 4 g (taps combined with 2 w both ways), so r
            // and q are formed at four times their value: every FP32 operand
            // scaled by 4, which rounds alike; the sums are scaled back in FP64
            const float kf4 = 4.f * kf, ck4 = 4.f * ck, c1k4 = 4.f * c1k;
            float acc[2 + NJ];
#pragma unroll
            for (int q = 0; q < 2 + NJ; ++q) acc[q] = 0.f;
            // a warped g of zero (the quadtree rules' mask regions), below
            // 1e-10 in magnitude (outside the nonzero set the statistics
            // take) or NaN (logical(NaN) throws): the general kernel decides
            bool small = false;
            const int lane_i = fresh_lane();  // for this iteration's f reads
#pragma unroll
            for (int k = 0; k < PPT; ++k) {
                // no lane of this wavefront holds a pixel of slot k
                if (64 * (k * NW + wv) >= npix) continue;
                const bool valid = (vbits >> k) & 1u;
                float dx, dy;
                unpack_dd(ddr[k], dx, dy);
                const float g4 = split_gather4(sw, tile, tnx, dx, dy);
                const float fv = fwin[pixel_of(lane_i, k)];
                small |= valid && !(fabsf(g4) > 4.f * 1e-10f);  // |g| > 1e-10
                const float r = valid ? fmaf(-kf4, fv, g4) - ck4 : 0.f;
                const float qv = fmaf(kf4, fv, g4) - c1k4;
                acc[0] += r;
                acc[1] = fmaf(r, qv, acc[1]);
                const float u = fxr[k] * r, v = fyr[k] * r;
                if (MODE == 6) {
                    acc[2] = fmaf(dx, u, acc[2]);
                    acc[3] = fmaf(dx, v, acc[3]);
                    acc[4] = fmaf(dy, u, acc[4]);
                    acc[5] = fmaf(dy, v, acc[5]);
                    acc[6] += u;
                    acc[7] += v;
                } else {
                    acc[2] += u;
                    acc[3] += v;
                }
            }
            int xv[1] = {quad && wave_any(small) ? 1 : 0};
            group_reduce<NW, 2 + NJ, 0, 1>(L, par, acc, nullptr, xv);
            if (xv[0] > 0) {
                state = 3;  // the general kernel (mask regions, statistics, NaN status)
                break;
            }
            // sum r, sum J r at 4 x, sum r q at 16 x their value
            const double s_r = 0.25 * (double)acc[0], s_rq = 0.0625 * (double)acc[1];
            // every window pixel is in the nonzero set of g
            const double n = npix;
            const double bf = bottomf;
            // statistics of g from the residual-form sums (residual_stats in
            // wave_common.h; T follows in the Gauss-Newton step)
            double dg, var;
            {
                const double k = kf;
                const double corr = (double)ck - (1.0 - k) * (double)cf;
                const double corr1 = (double)c1k - (1.0 + k) * (double)cf;
                const double F1c = L.F1c, F2c = L.F2c;
                const double S1 = s_r + k * F1c + n * corr;
                const double S2 = s_rq + k * k * F2c + corr1 * (S1 - k * F1c) +
                                  corr * (S1 + k * F1c) - n * corr * corr1;
                dg = S1 * inv_n;
                var = S2 - S1 * dg;
            }
            if (!(n >= 2) || !(bf > 0) || !(var > 0) || !isfinite(bf) || !isfinite(var) ||
                !isfinite(dg)) {
                state = 3;  // degenerate statistics: the reference's FP64 residual
                break;
            }
            const double sv = sqrt(var);
            kap = sv * inv_bf;
            // steepest-descent vector b = FJ - (bottomf/bottomg) (T - (gbar-c) SJ),
            // norm test (with its guard band), Gauss-Newton step and
            // inverse-compositional update; returns 0 to go on, 1 when the
            // loop ends, 2 when a test fell inside the guard band
            auto gn_step = [&]() -> int {
                double b[6] = {0, 0, 0, 0, 0, 0};
                {
                    // T_q = sum J_q r + kf FJc_q + corr SJ_q (residual_stats)
                    const double k = kf;
                    const double corr = (double)ck - (1.0 - k) * (double)cf;
                    const double ratio = bf / sv;
#pragma unroll
                    for (int q = 0; q < NJ; ++q) {
                        const double Tq = 0.25 * (double)acc[2 + q] + k * L.FJc[q] + corr * L.SJ[q];
                        b[q] = L.FJ[q] - ratio * (Tq - dg * L.SJ[q]);
                    }
                }
                double nb2 = 0.0;
                if (MODE == 6) {
#pragma unroll
                    for (int q = 0; q < 6; ++q) nb2 += b[q] * b[q];
                } else {
                    // :379 tempb = 2 b / bottomf^2 + mu (P - UOld - vdual)
                    const double s2 = 2.0 / (bottomf * bottomf);
                    b[0] = b[0] * s2 + a.mu * ((P[4] - uo_u) - vd_u);
                    b[1] = b[1] * s2 + a.mu * ((P[5] - uo_v) - vd_v);
                    nb2 = b[0] * b[0] + b[1] * b[1];
                }
                abs2 = nb2;
                if (it == 1) {
                    init_ok = nb2 > tol2;  // normOfWNewInit > tol
                    inv_init2 = 1.0 / nb2;
                }
                rel2 = init_ok ? nb2 * inv_init2 : 0.0;
                // near_tol against the band limits set once per launch
                if ((abs2 > a.abs_lo2 && abs2 < a.abs_hi2) ||
                    (init_ok && rel2 > a.rel_lo2 && rel2 < a.rel_hi2))
                    return 2;
                if (rel2 < tol2 || abs2 < tol2) return 1;
                double d[6] = {0, 0, 0, 0, 0, 0};
                if (MODE == 6) {
                    solve_chol6(L.Lp, L.Ld, b, d);
                } else {
                    d[4] = -(L.H2i[0] * b[0] + L.H2i[2] * b[1]);
                    d[5] = -(L.H2i[1] * b[0] + L.H2i[3] * b[1]);
                }
                return ic_compose(P, d) ? 0 : 1;
            };
            int go;
            if constexpr (NW >= 8) {
                // with eight or more wavefronts on one CU the redundant FP64
                // step would queue behind itself on every SIMD: wavefront 0
                // takes it and hands the new state to the others
                go = 0;
                if (wv == 0) {
                    go = gn_step();
                    if (lane == 0) {
#pragma unroll
                        for (int q = 0; q < 6; ++q) L.Pn[q] = P[q];
                        L.rel2 = rel2;
                        L.abs2 = abs2;
                        L.inv_init2 = inv_init2;
                        L.init_ok = init_ok;
                        L.stop = go;
                    }
                }
                __syncthreads();
                if (wv != 0) {
#pragma unroll
                    for (int q = 0; q < 6; ++q) P[q] = L.Pn[q];
                    rel2 = L.rel2;
                    abs2 = L.abs2;
                    inv_init2 = L.inv_init2;
                    init_ok = L.init_ok;
                    go = L.stop;
                }
                go = __builtin_amdgcn_readfirstlane(go);
            } else {
                // up to four wavefronts (one per SIMD): every wavefront takes
                // the same step, with no barrier on the iteration's critical path
                go = gn_step();
            }
            if (go == 2) {
                state = 3;
                exact = true;
                break;
            }
            if (go == 1) {
                state = 1;
                break;
            }
        }

        // ---------------------------------------------------------- outputs
        if (state == 1) {
            int st = it;
            if (!(rel2 < tol2 || abs2 < tol2)) st = a.max_it + 1;
            if (isnan(rel2)) st = a.max_it + 1;
            if (quad && sabs < 1e-6) st = a.max_it + 3;
            status = st;
        }
        if (tid == 0) {
            if (state == 3) {
                a.status[idx] = exact ? ST_EXACT : ST_DEFER;
                defer_append(a, idx, gld(a.wx, (size_t)idx), MODE == 2 ? gld(a.wy, (size_t)idx)
                                                                     : gld(a.wx, (size_t)idx));
                if (exact) {
                    if (a.exact_count) atomicAdd(a.exact_count, 1);
                } else if (a.handed) {
                    a.handed[idx] = 1;  // the schedule moves it to the general kernel
                }
            } else {
                double Pout[6];
                for (int k = 0; k < 6; ++k) Pout[k] = P[k];
                if (status == ST_ERROR)
                    for (int k = 0; k < 6; ++k) Pout[k] = NAN;
                if (status == ST_MASKED && MODE == 6) {
                    Pout[0] = Pout[1] = Pout[2] = Pout[3] = 0.0;
                    Pout[4] = gld(a.U0, (size_t)(2 * idx));
                    Pout[5] = gld(a.U0, (size_t)(2 * idx + 1));
                }
                a.U[2 * idx] = Pout[4];
                a.U[2 * idx + 1] = Pout[5];
                if (MODE == 6)
                    for (int k = 0; k < 4; ++k) a.F[4 * idx + k] = Pout[k];
                a.status[idx] = status;
            }
        }
        // the reference's H of zeros for these codes, one entry per lane
        if (MODE == 6 && state == 2 && (status == ST_ERROR || status == ST_MASKED) && tid < 21)
            a.H21[idx + (long)tid * a.N] = 0.0;
        __syncthreads();
        if (!a.work_counter) break;
    }
}

template <int PPT, int MODE, int NW>
hipError_t launch_lean(const IcgnArgs &a, int grid, hipStream_t s)
{
    // the g tile, then f of every pixel slot
    const size_t lds = sizeof(float) * ((size_t)a.tile_cap + 64 * PPT * NW);
    hipLaunchKernelGGL((icgn_lean_kernel<PPT, MODE, NW>), dim3(grid), dim3(64 * NW), lds, s, a);
    return hipGetLastError();
}

// the pixel-slot counts of icgn_wave.hip, per number of wavefronts
template <int MODE, int NW>
hipError_t launch_lean_ppt(const IcgnArgs &a, int ppt, int grid, hipStream_t s)
{
    if constexpr (NW == 1) {
        switch (ppt) {
        case 5: return launch_lean<5, MODE, 1>(a, grid, s);
        case 6: return launch_lean<6, MODE, 1>(a, grid, s);
        case 7: return launch_lean<7, MODE, 1>(a, grid, s);
        case 9: return launch_lean<9, MODE, 1>(a, grid, s);
        case 18: return launch_lean<18, MODE, 1>(a, grid, s);
        }
    } else if constexpr (NW == 2) {
        switch (ppt) {
        case 3: return launch_lean<3, MODE, 2>(a, grid, s);
        case 4: return launch_lean<4, MODE, 2>(a, grid, s);
        case 5: return launch_lean<5, MODE, 2>(a, grid, s);
        case 7: return launch_lean<7, MODE, 2>(a, grid, s);
        case 9: return launch_lean<9, MODE, 2>(a, grid, s);
        }
    } else if constexpr (NW == 4) {
        switch (ppt) {
        case 2: return launch_lean<2, MODE, 4>(a, grid, s);
        case 3: return launch_lean<3, MODE, 4>(a, grid, s);
        case 5: return launch_lean<5, MODE, 4>(a, grid, s);
        case 7: return launch_lean<7, MODE, 4>(a, grid, s);
        case 9: return launch_lean<9, MODE, 4>(a, grid, s);
        }
    } else if constexpr (NW == 8) {
        if (ppt == 9) return launch_lean<9, MODE, 8>(a, grid, s);
    } else if constexpr (NW == 16) {
        if (ppt == 5) return launch_lean<5, MODE, 16>(a, grid, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace

int icgn_lean_ppt(int npix_cap, int nw)
{
    // up to four wavefronts: the slot counts of the general kernel; eight and
    // sixteen wavefronts (windows of up to 4608 and 5120 pixels, the 64-px
    // windows): nine and five slots per lane
    if (nw <= 4) return icgn_wave_ppt(npix_cap, nw);
    if (nw == 8) return 64 * 8 * 9 >= npix_cap ? 9 : 0;
    if (nw == 16) return 64 * 16 * 5 >= npix_cap ? 5 : 0;
    return 0;
}

hipError_t icgn_lean_launch(const IcgnArgs &a, int mode, int ppt, int nw, int grid, hipStream_t s)
{
    if (mode == 6) {
        if (nw == 1) return launch_lean_ppt<6, 1>(a, ppt, grid, s);
        if (nw == 2) return launch_lean_ppt<6, 2>(a, ppt, grid, s);
        if (nw == 4) return launch_lean_ppt<6, 4>(a, ppt, grid, s);
        if (nw == 8) return launch_lean_ppt<6, 8>(a, ppt, grid, s);
        if (nw == 16) return launch_lean_ppt<6, 16>(a, ppt, grid, s);
    } else {
        if (nw == 1) return launch_lean_ppt<2, 1>(a, ppt, grid, s);
        if (nw == 2) return launch_lean_ppt<2, 2>(a, ppt, grid, s);
        if (nw == 4) return launch_lean_ppt<2, 4>(a, ppt, grid, s);
        if (nw == 8) return launch_lean_ppt<2, 8>(a, ppt, grid, s);
        if (nw == 16) return launch_lean_ppt<2, 16>(a, ppt, grid, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace staq
