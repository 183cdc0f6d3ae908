// Wave-per-subset IC-GN kernel for gfx950 (the fast path of the batched local
// step).  Same semantics as icgn.hip (funICGNQuadtree.m / funICGN.m /
// funICGN_Subpb1.m); differs only in how the work is laid out:
//
//   * one 64-lane wavefront fits one subset; a workgroup is one wavefront, so
//     every barrier is a single-wave s_barrier;
//   * each lane keeps its pixels' df/dx, df/dy and g in registers (PPT pixels
//     per lane, pixel p = lane + 64 k of the column-major window);
//   * the per-subset control state (warp P, norms, statistics) is computed
//     redundantly by all 64 lanes in FP64 registers, so the Gauss-Newton step
//     needs no serial single-thread section and no LDS round trip;
//   * sums are FP32 per lane and reduced with a butterfly (every lane ends
//     with the total); the inverse Hessian (FP64) sits in LDS and is read with
//     broadcast loads;
//   * bwselect regions are flood-filled bit-parallel: window column j is a
//     64-bit word held by lane j, one propagation step is a few shifts and two
//     cross-lane shuffles.
//
// Windows that do not fit 64 * PPT pixels after the mask-driven growth are
// deferred (status -2) to the workgroup kernel of icgn.hip.
#include "wave_common.h"
#include "flood.h"

namespace staq {

namespace {

struct WaveLds;

struct WaveLds {
    double Lp[21];     // packed Cholesky factor of H (column-major lower triangle)
    double Ld[6];      // 1 / diag(Lp)
    double A[36];      // LU factors when H is not positive definite
    double work[36];
    int piv[6], kind;  // kind 0: Cholesky in Lp, 1: LU in A
    double H2i[4];     // 2-DOF: inverse of the damped 2x2 system
    double FJ[6], SJ[6];
    // the same sums about cf = (float) meanf (FJc = sum J (f - cf),
    // F1c = sum (f - cf), F2c = sum (f - cf)^2), for the residual form
    double FJc[6], F1c, F2c;
    unsigned long long colw[64];  // scratch column words (atomic OR target)
    unsigned long long tmw[64];   // window-mask columns
    unsigned long long sel0w[64]; // setup region (4-connected mask region)
    // exchange between the wavefronts of one subset (double-buffered)
    float xf[2][4][24];
    double xd[2][4][16];
    int xi[2][4][8];
    unsigned long long bal[64];   // per pixel slot ballots of nonzero g
};

// Solve H d = -b with the factor in LDS (all lanes read it with broadcast loads).
__device__ __forceinline__ void solve6_lds(const WaveLds &L, const double (&b)[6], double (&d)[6])
{
    double y[6];
    if (L.kind == 0) {
        solve_chol6(L.Lp, L.Ld, b, d);
    } else {
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            double s = -b[L.piv[i]];
            for (int k = 0; k < i; ++k) s -= L.A[i + k * 6] * y[k];
            y[i] = s;
        }
        for (int i = 5; i >= 0; --i) {
            double s = y[i];
            for (int k = i + 1; k < 6; ++k) s -= L.A[i + k * 6] * d[k];
            d[i] = s / L.A[i + i * 6];
        }
    }
}

template <int PPT, int MODE, int NW>
__global__ __launch_bounds__(64 * NW) void icgn_wave_kernel(IcgnArgs a)
{
    constexpr int NJ = MODE;
    constexpr int CAP = 64 * PPT * NW;
    constexpr bool BRANCH_FREE = PPT <= 9;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ WaveLds L;
    __shared__ long item_sh;
    float *tile = reinterpret_cast<float *>(smem);
    // f of every pixel slot on the set-up's region (0 outside it), read each
    // iteration by the residual form of the sums; after the g tile
    float *fwin = tile + a.tile_cap;
    const int tid = threadIdx.x;
    const int lane = tid & 63;   // lane within the wavefront
    const int wv = tid >> 6;     // wavefront of the subset
    const bool quad = !(MODE == 6 && a.plain);
    int par = 0;                 // exchange buffer parity
    // pixel slot k of this lane: slots are interleaved over the wavefronts
    auto pixel = [&](int k) { return lane + 64 * (k * NW + wv); };

    for (long pass = 0;; ++pass) {
        if (tid == 0)
            item_sh = a.work_counter ? a.item0 + atomicAdd(a.work_counter, 1)
                                     : (pass == 0 ? a.item0 + blockIdx.x : a.item0 + a.n_items);
        __syncthreads();
        const long item = item_sh;
        if (item >= a.item0 + a.n_items) break;
        const long idx = a.order ? gld(a.order, (size_t)(item)) : item;
        const int slot = a.slot ? gld(a.slot, (size_t)(idx)) : 0;
        const FrameDev fr = load_frame(a.frames, slot);
        const int Hd = fr.H - 6;

        double x0 = gld(a.coords, (size_t)(idx)), y0 = gld(a.coords, (size_t)(idx + a.N));
        if (!a.plain) {
            x0 = round(x0);  // LocalICGNQuadtree.m:69
            y0 = round(y0);
        }
        int wx = gld(a.wx, (size_t)(idx));
        int wy = MODE == 2 ? gld(a.wy, (size_t)(idx)) : wx;
        const int w0 = wx;
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
        auto win_fits = [&](int wxx, int wyy) {
            const double xa = x0 - wxx * 0.5, xb = x0 + wxx * 0.5, ya = y0 - wyy * 0.5,
                         yb = y0 + wyy * 0.5;
            if (xa != floor(xa) || ya != floor(ya) || xb != floor(xb) || yb != floor(yb)) return false;
            return xa >= 4 && ya >= 4 && xb <= fr.H - 3 && yb <= fr.W - 3;
        };
        if (!win_fits(wx, wy)) {
            state = 2;
            status = ST_ERROR;
        } else if ((wx + 1) * (wy + 1) > CAP || wx + 1 > 64 || wy + 1 > 64) {
            state = 3;
        }

        // ------------------------------------------------------------ setup
        int nx = wx + 1, ny = wy + 1, npix = nx * ny;
        int xa = (int)x0 - wx / 2, ya = (int)y0 - wy / 2;
        int count = 0;
        if (state == 0) {
            // masked-pixel count (funICGNQuadtree.m:61 / funICGN.m:54)
            int cnt = 0;
            const float rnx = 1.0f / (float)nx;
#pragma unroll
            for (int k = 0; k < PPT; ++k) {
                const int p = pixel(k);
                if (p < npix) {
                    const int j = (int)(((float)p + 0.5f) * rnx), i = p - j * nx;
                    const int x = xa + i, y = ya + j;
                    if (MODE == 6 && a.plain)
                        cnt += gld(fr.dfdx, (size_t)((x - 4) + (y - 4) * Hd)) == 0.f;
                    else
                        cnt += fidx(fr.f, fr.H, x, y) * fidx(fr.mask, fr.H, x, y) == 0.f;
                }
            }
            {
                int ci[1] = {(int)wave_sum_dpp((float)cnt)};
                group_reduce<NW, 0, 0, 1>(L, par, nullptr, nullptr, ci);
                count = ci[0];
            }
            const double n0 = (double)nx * ny;
            const double cap = (MODE == 6 && a.plain) ? 0.5 * n0 : 0.4 * n0;
            if (!(count < cap)) {
                state = 2;
                status = ST_MASKED;
            } else {
                bool grow;
                int gx = wx, gy = wy;
                if (MODE == 6 && a.plain) {
                    grow = count > 0.1 * n0;
                    if (grow) gx = gy = (int)round(sqrt(2.0) * wx);
                } else {
                    grow = count > 0;
                    if (grow) {
                        gx = 2 * (int)ceil(sqrt(n0 / (n0 - count)) * wx / 2.0);
                        if (MODE == 6) {
                            gy = gx;
                        } else {
                            const double n1 = (double)(gx + 1) * (wy + 1);
                            gy = 2 * (int)ceil(sqrt(n1 / (n1 - count)) * wy / 2.0);
                        }
                    }
                }
                if (grow) {
                    if (!win_fits(gx, gy)) {
                        state = 2;
                        status = ST_ERROR;
                    } else if ((gx + 1) * (gy + 1) > CAP || gx + 1 > 64 || gy + 1 > 64) {
                        state = 3;
                    } else {
                        wx = gx;
                        wy = gy;
                        nx = wx + 1;
                        ny = wy + 1;
                        npix = nx * ny;
                        xa = (int)x0 - wx / 2;
                        ya = (int)y0 - wy / 2;
                    }
                }
            }
        }
        const bool masked_f = quad && count > 0;
        const float rnx = 1.0f / (float)nx;
        const float hx = wx * 0.5f, hy = wy * 0.5f;
        const int seed_c = (wx + 1) / 2, seed_r = (wy + 1) / 2;  // bwselect(BW, c, r): c col, r row

        if (state == 0 && masked_f) {
            // window mask columns and the 4-connected region through the seed (:83-88)
            L.colw[lane] = 0ull;
            __syncthreads();
#pragma unroll
            for (int k = 0; k < PPT; ++k) {
                const int p = pixel(k);
                if (p < npix) {
                    const int j = (int)(((float)p + 0.5f) * rnx), i = p - j * nx;
                    if (fidx(fr.mask, fr.H, xa + i, ya + j) != 0.f) atomicOr(&L.colw[j], 1ull << i);
                }
            }
            __syncthreads();
            const unsigned long long tm = L.colw[lane];
            const unsigned long long s0w = wave_flood(tm, ny, seed_r, seed_c, 4);
            __syncthreads();
            if (wv == 0) {
                L.tmw[lane] = tm;
                L.sel0w[lane] = s0w;
            }
            __syncthreads();
        }

        // per-pixel df/dx, df/dy in registers, s0bits = pixels inside the
        // setup region; f statistics and steepest-descent sums in one pass
        // about the shift c0 = f at the centre
        float fxr[PPT], fyr[PPT];
        unsigned s0bits = ~0u;
        // convergence bookkeeping on squared norms (funICGNQuadtree.m:373-386):
        // rel2 = (|b| / |b_1|)^2 or 0, abs2 = |b|^2; a square root is taken
        // nowhere on the iteration's critical path
        double meanf = 0, bottomf = 0, sabs = 0, rel2 = 1, abs2 = 1, inv_init2 = 0;
        bool init_ok = false;
        bool rform_f = false;  // the set-up allows the residual form (no masked pixel)
        const double tol2 = a.tol * a.tol;
        if (state == 0) {
            const float c0 = fidx(fr.f, fr.H, xa + nx / 2, ya + ny / 2) *
                             (quad ? fidx(fr.mask, fr.H, xa + nx / 2, ya + ny / 2) : 1.f);
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
                fxr[k] = fyr[k] = 0.f;
                if (p < npix) {
                    const int j = (int)(((float)p + 0.5f) * rnx), i = p - j * nx;
                    const int x = xa + i, y = ya + j;
                    const size_t gi = (size_t)((x - 4) + (y - 4) * Hd);
                    fxr[k] = gld(fr.dfdx, gi);
                    fyr[k] = gld(fr.dfdy, gi);
                    float f = fidx(fr.f, fr.H, x, y);
                    if (quad) f *= fidx(fr.mask, fr.H, x, y);
                    const bool s = !masked_f || ((L.sel0w[j] >> i) & 1ull);
                    if (!s) s0bits &= ~(1u << k);
                    const float fx = s ? fxr[k] : 0.f, fy = s ? fyr[k] : 0.f;
                    if (!s) f = 0.f;
                    fwin[p] = f;
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
            double sd[4 + 2 * NJ];
            sd[0] = s_n;
            sd[1] = s_1;
            sd[2] = s_2;
            sd[3] = s_abs;
#pragma unroll
            for (int q = 0; q < NJ; ++q) {
                sd[4 + q] = sj[q];
                sd[4 + NJ + q] = sjf[q];
            }
            group_reduce<NW, NH, 4 + 2 * NJ, 0>(L, par, hs, sd, nullptr);
            const double n = sd[0];
            const double S1 = sd[1], S2 = sd[2];
            sabs = sd[3];
            meanf = n > 0 ? (double)c0 + S1 / n : NAN;
            bottomf = n >= 2 ? sqrt(fmax(S2 - S1 * S1 / n, 0.0)) : (n == 1 ? 0.0 : NAN);
            // the residual form of the sums needs every pixel in f's nonzero set
            rform_f = !masked_f && n == (double)npix;
            const double dm = (double)c0 - meanf;
            const double e0 = (double)c0 - (double)(isfinite(meanf) ? (float)meanf : 0.f);
#pragma unroll
            for (int q = 0; q < NJ; ++q) {
                const double SJq = sd[4 + q];
                const double SJFq = sd[4 + NJ + q];
                if (tid == 0) {
                    L.SJ[q] = SJq;
                    L.FJ[q] = SJFq + dm * SJq;   // sum J (f - meanf)
                    L.FJc[q] = SJFq + e0 * SJq;  // sum J (f - cf)
                }
            }
            if (tid == 0) {
                L.F1c = S1 + n * e0;
                L.F2c = S2 + 2.0 * e0 * S1 + n * e0 * e0;
            }
            if (MODE == 6) {
                double h[21];
#pragma unroll
                for (int q = 0; q < 21; ++q) h[q] = hs[q];
                if (tid == 0)
                    for (int q = 0; q < 21; ++q) a.H21[idx + (long)q * a.N] = h[q];
                const bool spd = chol6(h);
                if (a.guard2 > 0) {
                    // an ill-conditioned or indefinite H: the FP64 re-solve
                    double dg6[6];
#pragma unroll
                    for (int c = 0; c < 6; ++c) dg6[c] = h[c * 6 - (c * (c - 1)) / 2];
                    if (!spd || pivots_spread(dg6)) {
                        state = 3;
                        exact = true;
                    }
                }
                if (tid == 0) {
                    if (spd) {
                        for (int q = 0; q < 21; ++q) L.Lp[q] = h[q];
                        for (int c = 0; c < 6; ++c) L.Ld[c] = 1.0 / h[c * 6 - (c * (c - 1)) / 2];
                        L.kind = 0;
                    } else {
                        int t = 0;
                        for (int c = 0; c < 6; ++c)
                            for (int r = c; r < 6; ++r) {
                                L.A[r + c * 6] = L.A[c + r * 6] = (double)hs[t];
                                ++t;
                            }
                        int kind;
                        factor_lds<6>(L.A, L.work, L.piv, &kind);
                        if (kind == 0) {  // the trial Cholesky in LDS succeeded after all
                            for (int c = 0, t2 = 0; c < 6; ++c)
                                for (int r = c; r < 6; ++r) L.Lp[t2++] = L.A[r + c * 6];
                            for (int c = 0; c < 6; ++c) L.Ld[c] = 1.0 / L.A[c + c * 6];
                        }
                        L.kind = kind;
                    }
                }
            } else {
                set_h2(L, hs[0], hs[1], hs[2], bottomf, a.mu);
            }
            __syncthreads();
        }

        // ---------------------------------------------------------- iterate
        int it = 0, sum_bw2 = 0;
        bool fsel = false;  // 2-DOF: f side taken from the last g region
        unsigned selbits = ~0u;  // pixels of the current g region (bit k = pixel lane + 64 k)
        // the g region is a function of the zero pattern of g alone: keep the
        // pattern it was computed from (nzprev) and reuse it while unchanged;
        // spec_next = the last iteration was masked, so the next pass applies it
        unsigned nzprev = 0u;
        bool region_valid = false, spec_next = false, region_new_prev = false;
        // the region's column word (this lane's column) and whether the
        // region is the whole nonzero set of g, the usual case
        unsigned long long regw = 0ull;
        bool region_is_nz = false;
        // g tile resident in LDS: [tu0, tu0+tnx) x [tv0, tv0+tny) (1-based)
        int tu0 = 0, tv0 = 0, tnx = 0, tny = 0;
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
            if (MODE == 6 && quad && it > 1 && count > 0 && region_new_prev) {
                // funICGNQuadtree.m:139-145 (see icgn.hip); the size depends
                // on the last g region only, so it is re-derived when that changed
                region_new_prev = false;
                const double wn =
                    2.0 * ceil(sqrt((double)(w0 + 1) * (w0 + 1) / (double)sum_bw2) * wx / 2.0);
                if (wn != (double)wx) {
                    if (!isfinite(wn) || x0 - wn / 2 < 1 || x0 + wn / 2 > fr.H ||
                        y0 - wn / 2 < 1 || y0 + wn / 2 > fr.W) {
                        state = 2;
                        status = ST_ERROR;
                        break;
                    }
                    const double h = wn / 2;
                    bool out = false;
                    for (int q = 0; q < 4; ++q) {
                        const double dx = (q & 1) ? h : -h, dy = (q & 2) ? h : -h;
                        const double u = (1 + P[0]) * dx + P[2] * dy + (x0 + P[4]);
                        const double v = P[1] * dx + (1 + P[3]) * dy + (y0 + P[5]);
                        if (u < 3 || u > fr.H - 2 || v < 3 || v > fr.W - 2) out = true;
                    }
                    if (out) {
                        rel2 = 1e12;  // normOfWNew = 1e6 (:156)
                        state = 1;
                    } else {
                        state = 2;
                        status = ST_ERROR;
                    }
                    break;
                }
            }
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
            // the warped window's extremes are at its corners (funICGNQuadtree.m:149-157),
            // computed alike in every lane (no cross-lane step on the
            // iteration's critical path)
            double umin, umax, vmin, vmax;
            bool out;
            {
                double u[4], v[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const double dx = (q & 1) ? hx : -hx, dy = (q & 2) ? hy : -hy;
                    u[q] = (1 + P[0]) * dx + P[2] * dy + (x0 + P[4]);
                    v[q] = P[1] * dx + (1 + P[3]) * dy + (y0 + P[5]);
                }
                umin = fmin(fmin(u[0], u[1]), fmin(u[2], u[3]));
                umax = fmax(fmax(u[0], u[1]), fmax(u[2], u[3]));
                vmin = fmin(fmin(v[0], v[1]), fmin(v[2], v[3]));
                vmax = fmax(fmax(v[0], v[1]), fmax(v[2], v[3]));
                bool o = false, edge = false;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    o |= (u[q] < 3) | (u[q] > fr.H - 2) | (v[q] < 3) | (v[q] > fr.W - 2);
                    edge |= near_edge(u[q], 3.0, fr.H - 2.0) | near_edge(v[q], 3.0, fr.W - 2.0);
                }
                out = o;
                if (edge && a.guard2 > 0) {
                    state = 3;  // the bounds test decided within FP32 reach: FP64 re-solve
                    exact = true;
                    break;
                }
            }
            if (out) {
                rel2 = 1e12;
                state = 1;
                break;
            }
            // support needed this iteration (taps floor-1..floor+2, one pixel
            // of slack for the FP32 local coordinates)
            const int nu0 = (int)floor(umin) - 2, nv0 = (int)floor(vmin) - 2;
            const int nu1 = (int)floor(umax) + 3, nv1 = (int)floor(vmax) + 3;
            bool use_tile = true;
            bool reload = !have_tile || nu0 < tu0 || nv0 < tv0 || nu1 >= tu0 + tnx ||
                          nv1 >= tv0 + tny;
            if (reload) {
                // reload with a margin so that the next iterations' small
                // updates stay inside the resident tile
                int m = 4;
                while (m > 0 && (nu1 - nu0 + 1 + 2 * m) * (nv1 - nv0 + 1 + 2 * m) > a.tile_cap) --m;
                tu0 = nu0 - m;
                tv0 = nv0 - m;
                tnx = nu1 - nu0 + 1 + 2 * m;
                tny = nv1 - nv0 + 1 + 2 * m;
                use_tile = tnx * tny <= a.tile_cap;
                have_tile = use_tile;
                if (!use_tile) {  // gather from HBM; origin as for a tile
                    tu0 = nu0;
                    tv0 = nv0;
                    tnx = nu1 - nu0 + 1;
                    tny = nv1 - nv0 + 1;
                }
            }
            // warp coordinates split into integer and FP32 fractional parts
            const SplitWarp sw = split_warp(P, x0, y0, tu0, tv0);
            if (use_tile && reload) {
                __syncthreads();  // last iteration's gathers are done with the old tile
                // flat over the tile, 8 independent loads in flight per lane
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
            // Keys bicubic of pixel (i, j) of the warped window, from the
            // resident tile or, if the warped support outgrew it, from HBM
            auto gather = [&](auto from_tile, int i, int j) {
                const float dx = (float)i - hx, dy = (float)j - hy;
                if constexpr (decltype(from_tile)::value) {
                    return split_gather(sw, tile, tnx, dx, dy);
                } else {
                    // the same taps and weights, read from HBM (clamped to the image)
                    const float su = fmaf(sw.p1, dx, fmaf(sw.p3, dy, sw.fu));
                    const float sv = fmaf(sw.p2, dx, fmaf(sw.p4, dy, sw.fv));
                    const float fsu = floorf(su), fsv = floorf(sv);
                    float wu[4], wv[4];
                    keys_w2(__builtin_amdgcn_fractf(su), wu);
                    keys_w2(__builtin_amdgcn_fractf(sv), wv);
                    const int iu = (int)((dx + sw.cu) + fsu), iv = (int)((dy + sw.cv) + fsv);
                    float g = 0.f;
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int y = min(max(tv0 + iv + q, 1), fr.W);
                        const float *col = fr.g + (size_t)(y - 1) * fr.H;
                        float s = 0.f;
#pragma unroll
                        for (int r = 0; r < 4; ++r) s = fmaf(wu[r], gld(col, (size_t)(min(max(tu0 + iu + r, 1), fr.H) - 1)), s);
                        g = fmaf(wv[q], s, g);
                    }
                    return 0.25f * g;  // the weights are 2 w
                }
            };
            auto gather_rt = [&](int i, int j) {
                return use_tile ? gather(std::true_type{}, i, j) : gather(std::false_type{}, i, j);
            };
            // Per-iteration sums: S1 = sum(g - c), S2 = sum((g - c)^2) over the
            // nonzero g, T_q = sum J_q (g - c) over the f-side pixels; counts
            // and the zero / NaN tests are wave votes.  When the last iteration
            // was masked, its g region is applied here already: the region
            // depends only on the zero pattern of g, which rarely changes from
            // one iteration to the next, so this pass is usually the final one.
            const bool gspec = spec_next;
            const unsigned fbits = (MODE == 2 && fsel) ? selbits : s0bits;
            // residual form of the sums (as icgn_lean.hip, same operations in
            // the same order): S1, S2 need the whole window as the nonzero set
            // (no mask region: rf), T the set-up's f side (rt; the 2-DOF
            // solver's rebuilt f side takes the plain sums)
            const bool rf = rform_f && !gspec && !(MODE == 2 && fsel);
            const bool rt = !(MODE == 2 && fsel);
            const float kf = (float)kap;
            const float ck = (float)((1.0 - (double)kf) * (double)cf);
            const float c1k = (float)((1.0 + (double)kf) * (double)cf);
            float gr[PPT];
            float acc[2 + NJ];
#pragma unroll
            for (int q = 0; q < 2 + NJ; ++q) acc[q] = 0.f;
            int n_nz = 0;
            unsigned nzb = 0u;
            bool any_zero = false, any_nan = false;
            // every lane runs every pixel slot (a lane past the window repeats
            // its last pixel and contributes nothing): no branch splits the
            // unrolled slots, so the gathers of several pixels are in flight
            // together
            // (up to 9 pixels per lane; with 18 the registers do not allow it
            // and each slot keeps its branch)
            auto fused_pass = [&](auto from_tile) {
#pragma unroll
                for (int k = 0; k < PPT; ++k) {
                    const int p = pixel(k);
                    const bool valid = p < npix;
                    bool nz = false;
                    if (BRANCH_FREE || valid) {
                    const int pc = valid ? p : npix - 1;
                    const int j = (int)(((float)pc + 0.5f) * rnx), i = pc - j * nx;
                    const float g = gather(from_tile, i, j);
                    gr[k] = valid ? g : 0.f;
                    any_zero |= valid && g == 0.f;
                    any_nan |= valid && isnan(g);
                    nzb |= (valid && g != 0.f) ? (1u << k) : 0u;
                    const float ge = (valid && (!gspec || ((selbits >> k) & 1u))) ? g : 0.f;
                    const float fv = fwin[p];
                    const float d = ge - cf;
                    const float r = fmaf(-kf, fv, ge) - ck;
                    const float qv = fmaf(kf, fv, ge) - c1k;
                    nz = valid && (!quad || fabsf(ge) > 1e-10f);
                    const float s1 = rf ? r : d, s2 = rf ? qv : d, t = rt ? r : d;
                    acc[0] += nz ? s1 : 0.f;
                    acc[1] = nz ? fmaf(s1, s2, acc[1]) : acc[1];
                    const bool s0 = valid && ((fbits >> k) & 1u);
                    const float u = s0 ? fxr[k] * t : 0.f, v = s0 ? fyr[k] * t : 0.f;
                    if (MODE == 6) {
                        const float dx = (float)i - hx, dy = (float)j - hy;
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
                    n_nz += (int)__popcll(__ballot(nz));
                }
            };
            if (use_tile)
                fused_pass(std::true_type{});
            else
                fused_pass(std::false_type{});
            int xv[4] = {n_nz, wave_any(any_zero) ? 1 : 0, wave_any(any_nan) ? 1 : 0,
                         wave_any(nzb != nzprev) ? 1 : 0};
            // one wavefront: the sums wait until it is known that they stand;
            // several: they travel with the votes, in one exchange
            if constexpr (NW == 1)
                group_reduce<NW, 0, 0, 4>(L, par, nullptr, nullptr, xv);
            else
                group_reduce<NW, 2 + NJ, 0, 4>(L, par, acc, nullptr, xv);
            n_nz = xv[0];
            if (quad && xv[2] > 0) {
                state = 2;
                status = ST_ERROR;  // logical(NaN) throws
                break;
            }
            const bool masked_it = quad && (count > 0 || xv[1] > 0);
            bool region_new = false;
            if (masked_it && !(region_valid && xv[3] == 0)) {
                // region of nonzero g through the seed: 8-connected for the
                // 6-DOF solver (funICGNQuadtree.m:171), 4-connected for
                // Subpb1 (funICGN_Subpb1.m:160); if that is the whole window,
                // the window mask's 8-connected region (:175-176 / :164-165)
                // column words from one ballot per pixel slot: slot k's ballot
                // is pixels 64k .. 64k+63 of the window in column-major order,
                // so column j is bits j*nx .. j*nx+nx-1 of that bit stream
                const unsigned long long full = nx >= 64 ? ~0ull : ((1ull << nx) - 1ull);
                unsigned long long colnz = 0ull;
                {
                    const int sb = lane * nx, wk = sb >> 6, off = sb & 63;
                    unsigned long long lo = 0ull, hi = 0ull;
                    if constexpr (NW == 1) {
#pragma unroll
                        for (int k = 0; k < PPT; ++k) {
                            const unsigned long long bk = __ballot((nzb >> k) & 1u);
                            lo = wk == k ? bk : lo;
                            hi = wk + 1 == k ? bk : hi;
                        }
                    } else {
#pragma unroll
                        for (int k = 0; k < PPT; ++k) {
                            const unsigned long long bk = __ballot((nzb >> k) & 1u);
                            if (lane == 0) L.bal[k * NW + wv] = bk;
                        }
                        if (lane == 0) L.bal[PPT * NW] = 0ull;
                        __syncthreads();
                        if (wk < PPT * NW) {
                            lo = L.bal[wk];
                            hi = L.bal[wk + 1];
                        }
                    }
                    colnz = off ? ((lo >> off) | (hi << (64 - off))) : lo;
                    colnz = lane < ny ? (colnz & full) : 0ull;
                }
                unsigned long long sw = wave_flood(colnz, ny, seed_r, seed_c, MODE == 6 ? 8 : 4);
                const bool all = !wave_any(lane < ny && sw != full);
                if (all) {
                    if (!masked_f) {
                        L.colw[lane] = 0ull;
                        __syncthreads();
#pragma unroll
                        for (int k = 0; k < PPT; ++k) {
                            const int p = pixel(k);
                            if (p < npix) {
                                const int j = (int)(((float)p + 0.5f) * rnx), i = p - j * nx;
                                if (fidx(fr.mask, fr.H, xa + i, ya + j) != 0.f)
                                    atomicOr(&L.colw[j], 1ull << i);
                            }
                        }
                        __syncthreads();
                        sw = wave_flood(L.colw[lane], ny, seed_r, seed_c, 8);
                    } else {
                        sw = wave_flood(L.tmw[lane], ny, seed_r, seed_c, 8);
                    }
                }
                if (lane >= ny) sw = 0ull;
                // (2-DOF treats every flood as a new region and rebuilds its
                // f side; the prediction below is for the 6-DOF solver)
                const bool changed = MODE == 2 || !region_valid || wave_any(sw != regw);
                region_is_nz = !wave_any(sw != colnz);
                regw = sw;
                __syncthreads();
                if (wv == 0) L.colw[lane] = sw;
                __syncthreads();
                if (changed) {
                selbits = 0u;
                sum_bw2 = 0;
#pragma unroll
                for (int k = 0; k < PPT; ++k) {
                    const int p = pixel(k);
                    bool in = false;
                    if (p < npix) {
                        const int j = (int)(((float)p + 0.5f) * rnx), i = p - j * nx;
                        in = (L.colw[j] >> i) & 1ull;
                    }
                    selbits |= in ? (1u << k) : 0u;
                    sum_bw2 += (int)__popcll(__ballot(in));
                }
                {
                    int sb2[1] = {sum_bw2};
                    group_reduce<NW, 0, 0, 1>(L, par, nullptr, nullptr, sb2);
                    sum_bw2 = sb2[0];
                }
                region_new = true;
                region_new_prev = true;
                if (MODE == 2) fsel = true;
                }
                nzprev = nzb;
                region_valid = true;
            }
            // The fused pass stands if it applied the region that holds now:
            // the last region (gspec) when that did not change, or (6-DOF) no
            // region when the region is the whole nonzero set of g (the pixels
            // outside it are zero anyway).
            bool stands;
            if (masked_it)
                stands = MODE == 6 ? (gspec ? !region_new : region_is_nz) : (gspec && !region_new);
            else
                stands = !gspec;
            const bool resum = !stands;
            spec_next = masked_it && (MODE == 2 || !region_is_nz);
            bool literal = false;
            double b[6] = {0, 0, 0, 0, 0, 0};
            if (!literal) {
                double n, S1, S2, T[6], SJv[6], FJv[6];
                double dg = 0, var = 0;
                bool have_stats = false;
#pragma unroll
                for (int q = 0; q < NJ; ++q) {
                    SJv[q] = L.SJ[q];
                    FJv[q] = L.FJ[q];
                }
                if (!resum) {
                    if constexpr (NW == 1) group_reduce<NW, 2 + NJ, 0, 0>(L, par, acc, nullptr, nullptr);
                    n = n_nz;
                    S1 = acc[0];
                    S2 = acc[1];
#pragma unroll
                    for (int q = 0; q < NJ; ++q) T[q] = acc[2 + q];
                    if (rf) {
                        // the residual form takes the nonzero set of g as the whole window
                        if (n == (double)npix) {
                            double Tr[NJ];
                            residual_stats<NJ>(n, acc, kf, ck, c1k, cf, L.F1c, L.F2c, L.FJc, L.SJ,
                                               dg, var, Tr);
#pragma unroll
                            for (int q = 0; q < NJ; ++q) T[q] = Tr[q];
                            have_stats = true;
                        } else {
                            literal = true;
                        }
                    } else if (rt) {
                        const double k = kf;
                        const double corr = (double)ck - (1.0 - k) * (double)cf;
#pragma unroll
                        for (int q = 0; q < NJ; ++q) T[q] = (double)acc[2 + q] + k * L.FJc[q] + corr * L.SJ[q];
                    }
                } else {
                    // the same sums with g on the current region (or on no
                    // region); Subpb1 also rebuilds f, grad f, H and the f
                    // statistics on a new region (funICGN_Subpb1.m:171-190),
                    // which then persist
                    const bool fre = MODE == 2 && region_new;
                    const unsigned fb = (MODE == 2 && fsel) ? selbits : s0bits;
                    float r[2 + NJ];
                    float q6[11];
#pragma unroll
                    for (int q = 0; q < 2 + NJ; ++q) r[q] = 0.f;
#pragma unroll
                    for (int q = 0; q < 11; ++q) q6[q] = 0.f;
                    int nr = 0;
#pragma unroll
                    for (int k = 0; k < PPT; ++k) {
                        const int p = pixel(k);
                        bool nz = false;
                        if (p < npix) {
                            const bool sg = !masked_it || ((selbits >> k) & 1u);
                            const float g = sg ? gr[k] : 0.f;
                            const float d = g - cf;
                            nz = fabsf(g) > 1e-10f;
                            r[0] += nz ? d : 0.f;
                            r[1] += nz ? d * d : 0.f;
                            const bool sf = (fb >> k) & 1u;
                            const float fx = sf ? fxr[k] : 0.f, fy = sf ? fyr[k] : 0.f;
                            if (MODE == 6) {
                                const int j = (int)(((float)p + 0.5f) * rnx), i = p - j * nx;
                                const float dx = (float)i - hx, dy = (float)j - hy;
                                const float u = fx * d, v = fy * d;
                                r[2] = fmaf(dx, u, r[2]);
                                r[3] = fmaf(dx, v, r[3]);
                                r[4] = fmaf(dy, u, r[4]);
                                r[5] = fmaf(dy, v, r[5]);
                                r[6] += u;
                                r[7] += v;
                            } else {
                                r[2] += fx * d;
                                r[3] += fy * d;
                                if (fre) {
                                    // f*mask on the new region, read again (the
                                    // region changes in few iterations)
                                    float tf = 0.f;
                                    if (sf) {
                                        const int j = (int)(((float)p + 0.5f) * rnx), i = p - j * nx;
                                        tf = fidx(fr.f, fr.H, xa + i, ya + j) * fidx(fr.mask, fr.H, xa + i, ya + j);
                                    }
                                    const float df = tf - cf;
                                    const bool fnz = fabsf(tf) > 1e-10f;
                                    q6[0] += fnz ? 1.f : 0.f;
                                    q6[1] += fnz ? df : 0.f;
                                    q6[2] += fnz ? df * df : 0.f;
                                    q6[3] += fabsf(tf);
                                    q6[4] += fx * fx;
                                    q6[5] += fx * fy;
                                    q6[6] += fy * fy;
                                    q6[7] += fx;
                                    q6[8] += fy;
                                    q6[9] += fx * df;
                                    q6[10] += fy * df;
                                }
                            }
                        }
                        nr += (int)__popcll(__ballot(nz));
                    }
                    {
                        int ni[1] = {nr};
                        group_reduce<NW, 2 + NJ, 0, 1>(L, par, r, nullptr, ni);
                        nr = ni[0];
                    }
                    n = nr;
                    S1 = r[0];
                    S2 = r[1];
#pragma unroll
                    for (int q = 0; q < NJ; ++q) T[q] = r[2 + q];
                    if (fre) {
                        double fd[8] = {q6[0], q6[1], q6[2], q6[3], q6[7], q6[8], q6[9], q6[10]};
                        float fh[3] = {q6[4], q6[5], q6[6]};
                        group_reduce<NW, 3, 8, 0>(L, par, fh, fd, nullptr);
                        const double nf = fd[0];
                        const double F1 = fd[1];
                        const double F2 = fd[2];
                        sabs = fd[3];
                        const double h11 = fh[0], h12 = fh[1], h22 = fh[2];
                        meanf = nf > 0 ? (double)cf + F1 / nf : NAN;
                        bottomf = nf >= 2 ? sqrt(fmax(F2 - F1 * F1 / nf, 0.0)) : (nf == 1 ? 0.0 : NAN);
                        const double dmf = (double)cf - meanf;
                        for (int q = 0; q < 2; ++q) {
                            SJv[q] = fd[4 + q];
                            FJv[q] = fd[6 + q] + dmf * SJv[q];
                        }
                        __syncthreads();
                        set_h2(L, h11, h12, h22, bottomf, a.mu);
                        if (tid == 0)
                            for (int q = 0; q < 2; ++q) {
                                L.SJ[q] = SJv[q];
                                L.FJ[q] = FJv[q];
                            }
                        __syncthreads();
                    }
                }
                // gbar - c = S1/n, bottomg^2 = S2 - S1^2/n; b = FJ - (bf/bg) (T - (gbar-c) SJ)
                if (!have_stats) {
                    dg = S1 / n;
                    var = S2 - S1 * dg;
                }
                const double bf = bottomf;
                if (literal || !(n >= 2) || !(bf > 0) || !(var > 0) || !isfinite(bf) ||
                    !isfinite(var) || !isfinite(dg)) {
                    literal = true;
                } else {
                    const double sv = sqrt(var);
                    const double ratio = bf / sv;
                    kap = sv / bf;
#pragma unroll
                    for (int q = 0; q < NJ; ++q) b[q] = FJv[q] - ratio * (T[q] - dg * SJv[q]);
                }
            }
            if (literal) {
                // ---- degenerate statistics: the reference's residual, FP64
                const bool fsl = MODE == 2 && fsel;
                // g of pixel k under the current region
                auto gsel = [&](int i, int j) {
                    const float g = gather_rt(i, j);
                    return (masked_it && !((L.colw[j] >> i) & 1ull)) ? 0.f : g;
                };
                // pass A: counts and first moments (2-DOF with the f side on
                // the last region also rebuilds f, grad f and H, :171-190)
                double v[7] = {0, 0, 0, 0, 0, 0, 0};
#pragma unroll 3
                for (int k = 0; k < PPT; ++k) {
                    const int p = pixel(k);
                    if (p < npix) {
                        const int j = (int)(((float)p + 0.5f) * rnx), i = p - j * nx;
                        const double g = gsel(i, j);
                        const bool nz = !quad || fabs(g) > 1e-10;
                        v[0] += nz ? 1.0 : 0.0;
                        v[1] += nz ? g : 0.0;
                        if (fsl) {
                            const bool s = (L.colw[j] >> i) & 1ull;
                            const int x = xa + i, y = ya + j;
                            const size_t gi = (size_t)((x - 4) + (y - 4) * Hd);
                            const double fx = s ? (double)gld(fr.dfdx, gi) : 0.0;
                            const double fy = s ? (double)gld(fr.dfdy, gi) : 0.0;
                            const double fv = s ? (double)fidx(fr.f, fr.H, x, y) * (double)fidx(fr.mask, fr.H, x, y) : 0.0;
                            v[2] += fabs(fv) > 1e-10 ? 1.0 : 0.0;
                            v[3] += fabs(fv) > 1e-10 ? fv : 0.0;
                            v[4] += fx * fx;
                            v[5] += fx * fy;
                            v[6] += fy * fy;
                            // |f| for the featureless test
                            v[2] += 0.0;
                        }
                    }
                }
                group_reduce<NW, 0, 7, 0>(L, par, nullptr, v, nullptr);
                const double ng = v[0];
                const double mg = ng > 0 ? v[1] / ng : NAN;
                double mf = meanf, bfx = bottomf, nf = 0, h11 = 0, h12 = 0, h22 = 0;
                if (fsl) {
                    nf = v[2];
                    mf = nf > 0 ? v[3] / nf : NAN;
                    h11 = v[4];
                    h12 = v[5];
                    h22 = v[6];
                }
                // pass A2: second moments about the means
                double s2g = 0.0, s2f = 0.0, sa = 0.0;
#pragma unroll 3
                for (int k = 0; k < PPT; ++k) {
                    const int p = pixel(k);
                    if (p < npix) {
                        const int j = (int)(((float)p + 0.5f) * rnx), i = p - j * nx;
                        const double g = gsel(i, j);
                        if (!quad || fabs(g) > 1e-10) s2g += (g - mg) * (g - mg);
                        if (fsl && ((L.colw[j] >> i) & 1ull)) {
                            const double fv = (double)fidx(fr.f, fr.H, xa + i, ya + j) *
                                              (double)fidx(fr.mask, fr.H, xa + i, ya + j);
                            if (fabs(fv) > 1e-10) s2f += (fv - mf) * (fv - mf);
                            sa += fabs(fv);
                        }
                    }
                }
                {
                    double t3[3] = {s2g, s2f, sa};
                    group_reduce<NW, 0, 3, 0>(L, par, nullptr, t3, nullptr);
                    s2g = t3[0];
                    s2f = t3[1];
                    sa = t3[2];
                }
                const double bg = ng >= 2 ? sqrt(s2g) : (ng == 1 ? 0.0 : NAN);
                if (fsl) {
                    sabs = sa;
                    bfx = nf >= 2 ? sqrt(s2f) : (nf == 1 ? 0.0 : NAN);
                    meanf = mf;
                    bottomf = bfx;
                    __syncthreads();
                    set_h2(L, h11, h12, h22, bfx, a.mu);
                    __syncthreads();
                }
                // pass B: b = bottomf * sum(J r) over every pixel (:362-371)
                double vb[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll 3
                for (int k = 0; k < PPT; ++k) {
                    const int p = pixel(k);
                    if (p < npix) {
                        const int j = (int)(((float)p + 0.5f) * rnx), i = p - j * nx;
                        const int x = xa + i, y = ya + j;
                        double fv, fx, fy;
                        if (fsl) {
                            const bool s = (L.colw[j] >> i) & 1ull;
                            const size_t gi = (size_t)((x - 4) + (y - 4) * Hd);
                            fx = s ? (double)gld(fr.dfdx, gi) : 0.0;
                            fy = s ? (double)gld(fr.dfdy, gi) : 0.0;
                            fv = s ? (double)fidx(fr.f, fr.H, x, y) * (double)fidx(fr.mask, fr.H, x, y) : 0.0;
                        } else {
                            const bool s = !masked_f || ((L.sel0w[j] >> i) & 1ull);
                            const size_t gi = (size_t)((x - 4) + (y - 4) * Hd);
                            fx = s ? (double)gld(fr.dfdx, gi) : 0.0;
                            fy = s ? (double)gld(fr.dfdy, gi) : 0.0;
                            fv = fidx(fr.f, fr.H, x, y);
                            if (quad) fv *= (double)fidx(fr.mask, fr.H, x, y);
                            if (!s) fv = 0.0;
                        }
                        const double r = (fv - mf) / bfx - ((double)gsel(i, j) - mg) / bg;
                        if (MODE == 6) {
                            const double dx = i - (double)hx, dy = j - (double)hy;
                            vb[0] += dx * fx * r;
                            vb[1] += dx * fy * r;
                            vb[2] += dy * fx * r;
                            vb[3] += dy * fy * r;
                            vb[4] += fx * r;
                            vb[5] += fy * r;
                        } else {
                            vb[0] += fx * r;
                            vb[1] += fy * r;
                        }
                    }
                }
                group_reduce<NW, 0, NJ, 0>(L, par, nullptr, vb, nullptr);
#pragma unroll
                for (int q = 0; q < NJ; ++q) b[q] = bfx * vb[q];
                if (bfx > 0 && bg > 0 && isfinite(bg / bfx)) kap = bg / bfx;
            }
            // norm test, Gauss-Newton step, inverse-compositional update
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
            if (near_tol(abs2, tol2, a.guard2_abs) || (init_ok && near_tol(rel2, tol2, a.guard2))) {
                state = 3;  // the test is within FP32 reach of tol: FP64 re-solve
                exact = true;
                break;
            }
            if (rel2 < tol2 || abs2 < tol2) {
                state = 1;
                break;
            }
            double d[6] = {0, 0, 0, 0, 0, 0};
            if (MODE == 6) {
                solve6_lds(L, b, d);
            } else {
                d[4] = -(L.H2i[0] * b[0] + L.H2i[2] * b[1]);
                d[5] = -(L.H2i[1] * b[0] + L.H2i[3] * b[1]);
            }
            if (!ic_compose(P, d)) {
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
                if (exact && a.exact_count) atomicAdd(a.exact_count, 1);
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
                if (MODE == 6) {
                    for (int k = 0; k < 4; ++k) a.F[4 * idx + k] = Pout[k];
                    if (status == ST_ERROR || status == ST_MASKED)
                        for (int k = 0; k < 21; ++k) a.H21[idx + (long)k * a.N] = 0.0;
                }
                a.status[idx] = status;
            }
        }
        __syncthreads();
        if (!a.work_counter) break;
    }
}

template <int PPT, int MODE, int NW>
hipError_t launch_wave(const IcgnArgs &a, int grid, hipStream_t s)
{
    // the g tile, then f of every pixel slot
    const size_t lds = sizeof(float) * ((size_t)a.tile_cap + 64 * PPT * NW);
    hipLaunchKernelGGL((icgn_wave_kernel<PPT, MODE, NW>), dim3(grid), dim3(64 * NW), lds, s, a);
    return hipGetLastError();
}

// pixel slots per lane available for each number of wavefronts per subset
template <int MODE, int NW>
hipError_t launch_wave_ppt(const IcgnArgs &a, int ppt, int grid, hipStream_t s)
{
    if constexpr (NW == 1) {
        switch (ppt) {
        case 5: return launch_wave<5, MODE, 1>(a, grid, s);
        case 6: return launch_wave<6, MODE, 1>(a, grid, s);
        case 7: return launch_wave<7, MODE, 1>(a, grid, s);
        case 9: return launch_wave<9, MODE, 1>(a, grid, s);
        case 18: return launch_wave<18, MODE, 1>(a, grid, s);
        }
    } else if constexpr (NW == 2) {
        switch (ppt) {
        case 3: return launch_wave<3, MODE, 2>(a, grid, s);
        case 4: return launch_wave<4, MODE, 2>(a, grid, s);
        case 5: return launch_wave<5, MODE, 2>(a, grid, s);
        case 7: return launch_wave<7, MODE, 2>(a, grid, s);
        case 9: return launch_wave<9, MODE, 2>(a, grid, s);
        }
    } else {
        switch (ppt) {
        case 2: return launch_wave<2, MODE, 4>(a, grid, s);
        case 3: return launch_wave<3, MODE, 4>(a, grid, s);
        case 5: return launch_wave<5, MODE, 4>(a, grid, s);
        case 7: return launch_wave<7, MODE, 4>(a, grid, s);
        case 9: return launch_wave<9, MODE, 4>(a, grid, s);
        }
    }
    return hipErrorInvalidValue;
}

}  // namespace

int icgn_wave_ppt(int npix_cap, int nw)
{
    // smallest instantiated pixels-per-lane count that holds npix_cap pixels
    // over nw wavefronts
    static const int s1[] = {5, 6, 7, 9, 18}, s2[] = {3, 4, 5, 7, 9}, s4[] = {2, 3, 5, 7, 9};
    const int *set = nw == 1 ? s1 : (nw == 2 ? s2 : s4);
    const int n = 5;
    if (nw != 1 && nw != 2 && nw != 4) return 0;
    for (int i = 0; i < n; ++i)
        if (64 * nw * set[i] >= npix_cap) return set[i];
    return 0;  // too large for the wave kernel
}

hipError_t icgn_wave_launch(const IcgnArgs &a, int mode, int ppt, int nw, int grid, hipStream_t s)
{
    if (mode == 6) {
        if (nw == 1) return launch_wave_ppt<6, 1>(a, ppt, grid, s);
        if (nw == 2) return launch_wave_ppt<6, 2>(a, ppt, grid, s);
        if (nw == 4) return launch_wave_ppt<6, 4>(a, ppt, grid, s);
    } else {
        if (nw == 1) return launch_wave_ppt<2, 1>(a, ppt, grid, s);
        if (nw == 2) return launch_wave_ppt<2, 2>(a, ppt, grid, s);
        if (nw == 4) return launch_wave_ppt<2, 4>(a, ppt, grid, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace staq
