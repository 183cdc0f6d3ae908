// Bicubic gather (ba_interp2 drop-in), reference-image gradient, and the
// ZNSSD forward/backward (adjoint) kernels for gfx950.
#include "kernels.h"

namespace staq {

// ---------------------------------------------------------------------------
// ba_interp2.cpp:184-234: Keys bicubic, taps clamped to the array (:60-130).
// F is M x N (x P layers), X = column (dim-2) coordinate, Y = row (dim-1)
// coordinate, 1-based; out is nd x P.  One thread per sample point; the
// clamp-free interior case is the common one and reads 4 contiguous rows.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void ba_interp2_kernel(const T *__restrict__ F, long M, long N,
                                                         long P, const T *__restrict__ X,
                                                         const T *__restrict__ Y, long nd,
                                                         T *__restrict__ out)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= nd) return;
    const T x = X[i], y = Y[i];
    if (isnan(x) || isnan(y)) {
        for (long p = 0; p < P; ++p) out[i + p * nd] = T(NAN);
        return;
    }
    const T xf = floor(x), yf = floor(y);
    T wx[4], wy[4];
    keys_w(x - xf, wx);
    keys_w(y - yf, wy);
    // 0-based index of the floor pixel; out-of-range values are clamped, as
    // the reference does, after conversion to integer
    const double xfd = (double)xf - 1.0, yfd = (double)yf - 1.0;
    const long cx = (long)fmin(fmax(xfd, -4.0), (double)N + 4.0);
    const long cy = (long)fmin(fmax(yfd, -4.0), (double)M + 4.0);
    long col[4], row[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        col[k] = min(max(cx - 1 + k, 0L), N - 1);
        row[k] = min(max(cy - 1 + k, 0L), M - 1);
    }
    const long LO = M * N;
    for (long p = 0; p < P; ++p) {
        const T *Fp = F + p * LO;
        T acc = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            T s = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) s += wx[k] * Fp[row[j] + col[k] * M];
            acc += wy[j] * s;
        }
        out[i + p * nd] = acc;
    }
}

hipError_t ba_interp2_launch(const void *F, long M, long N, long P, const void *X, const void *Y,
                             long nd, void *out, bool single, hipStream_t s)
{
    if (nd <= 0) return hipSuccess;
    const unsigned grid = (unsigned)((nd + 255) / 256);
    if (single)
        hipLaunchKernelGGL(ba_interp2_kernel<float>, dim3(grid), dim3(256), 0, s, (const float *)F, M,
                           N, P, (const float *)X, (const float *)Y, nd, (float *)out);
    else
        hipLaunchKernelGGL(ba_interp2_kernel<double>, dim3(grid), dim3(256), 0, s, (const double *)F,
                           M, N, P, (const double *)X, (const double *)Y, nd, (double *)out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// func/funImgGradient.m:53-78: 7-tap central difference (imfilter correlation
// with [-1/60 3/20 -3/4 0 3/4 -3/20 1/60]), cropped by 3, times the mask.
// The taps are summed in FP64 in filter order with contraction off and the
// result rounded to FP32: funICGN.m:54 tests DfDx == 0, so flat image regions
// must give the same exact zeros (or the same rounding residue) as an FP64
// evaluation of the filter, not whatever a reassociated FP32 sum gives.
// Thread (x, y) of the (H-6) x (W-6) output; consecutive threads walk dim 1.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void img_gradient_kernel(const float *__restrict__ f,
                                                           const float *__restrict__ mask, int H,
                                                           int W, float *__restrict__ dfdx,
                                                           float *__restrict__ dfdy)
{
#pragma clang fp contract(off)
    const double h[7] = {-1.0 / 60, 3.0 / 20, -3.0 / 4, 0.0, 3.0 / 4, -3.0 / 20, 1.0 / 60};
    const int Hd = H - 6, Wd = W - 6;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;  // 0..Hd-1
    const int j = blockIdx.y;                             // 0..Wd-1
    if (i >= Hd || j >= Wd) return;
    const int x = i + 3, y = j + 3;  // 0-based pixel
    const float *col = f + (size_t)y * H;
    double sx = 0.0, sy = 0.0;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        sx += h[k] * (double)col[x + k - 3];
        sy += h[k] * (double)f[x + (size_t)(y + k - 3) * H];
    }
    const double m = mask[x + (size_t)y * H];
    dfdx[i + (size_t)j * Hd] = (float)(sx * m);
    dfdy[i + (size_t)j * Hd] = (float)(sy * m);
}

hipError_t img_gradient_launch(const float *f, const float *mask, int H, int W, float *dfdx,
                               float *dfdy, hipStream_t s)
{
    if (H < 7 || W < 7) return hipErrorInvalidValue;
    dim3 grid((unsigned)((H - 6 + 255) / 256), (unsigned)(W - 6));
    hipLaunchKernelGGL(img_gradient_kernel, grid, dim3(256), 0, s, f, mask, H, W, dfdx, dfdy);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// ZNSSD mismatch and its adjoint.
//   C_i = sum_p (fhat_p - ghat_p)^2 = 2 - 2 ZNCC_i,   fhat = (f - fbar)/||f - fbar||,
//   ghat likewise with g_p = bicubic(g)(u22_p, v22_p) and whole-window
//   statistics.  The vector-Jacobian product w.r.t. the warp parameters is
//   dC/dP = sum_p dC/dg_p (dg/du u_P + dg/dv v_P),
//   dC/dg_p = -(2/||g - gbar||) (fhat_p - ZNCC ghat_p),
//   u_P = [dx 0 dy 0 1 0], v_P = [0 dx 0 dy 0 1].
// One workgroup per subset.  Pass 1 gathers g and its two derivatives (Keys
// kernel and its derivative on the same 16 taps) into LDS and reduces the
// five moments; pass 2 (backward only) forms the gradient.  Sums are FP32
// per lane and FP64 across the block.
// ---------------------------------------------------------------------------


template <int BS>
__global__ __launch_bounds__(BS) void znssd_kernel(ZnssdArgs a)
{
    constexpr int NW = BS / kWave;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ double redd[NW * 6];
    __shared__ double stat[8];
    __shared__ int ok_sh;
    float *gv = reinterpret_cast<float *>(smem);
    float *gu = gv + a.npix_cap;
    float *gw = gu + a.npix_cap;
    const long idx = blockIdx.x;
    const int tid = threadIdx.x;
    const FrameDev fr = a.frames[a.slot ? a.slot[idx] : 0];
    const int w = a.winsize[idx], nx = w + 1, npix = nx * nx;
    const double x0 = round(a.coords[idx]), y0 = round(a.coords[idx + a.N]);
    const double *P = a.P + 6 * idx;
    if (tid == 0) {
        // same window rule as the IC-GN kernels (gradient array bounds)
        bool ok = npix <= a.npix_cap && x0 - w / 2 >= 4 && y0 - w / 2 >= 4 &&
                  x0 + w / 2 <= fr.H - 3 && y0 + w / 2 <= fr.W - 3;
        const double h = w * 0.5;
        for (int q = 0; q < 4 && ok; ++q) {
            const double dx = (q & 1) ? h : -h, dy = (q & 2) ? h : -h;
            const double u = (1 + P[0]) * dx + P[2] * dy + (x0 + P[4]);
            const double v = P[1] * dx + (1 + P[3]) * dy + (y0 + P[5]);
            if (!(u >= 3 && u <= fr.H - 2 && v >= 3 && v <= fr.W - 2)) ok = false;
        }
        ok_sh = ok;
    }
    __syncthreads();
    if (!ok_sh) {
        if (tid == 0) {
            if (a.cost) a.cost[idx] = 0.0;
            if (a.gradP)
                for (int k = 0; k < 6; ++k) a.gradP[6 * idx + k] = 0.0;
            if (a.status) a.status[idx] = 1;
        }
        return;
    }
    const int xa = (int)x0 - w / 2, ya = (int)y0 - w / 2;
    // local warp relative to an integer origin near the window keeps the FP32
    // coordinates small
    const int ou = (int)floor(x0 + P[4]), ov = (int)floor(y0 + P[5]);
    const float a11 = (float)(1 + P[0]), a12 = (float)P[2], a21 = (float)P[1], a22 = (float)(1 + P[3]);
    const float cu = (float)((x0 + P[4]) - ou), cv = (float)((y0 + P[5]) - ov);
    const float rcp = 1.0f / nx, h = w * 0.5f;
    double s[5] = {0, 0, 0, 0, 0};  // sum f, f^2, g, g^2, f g
    for (int p = tid; p < npix; p += BS) {
        const int j = (int)(((float)p + 0.5f) * rcp), i = p - j * nx;
        const float dx = (float)i - h, dy = (float)j - h;
        const float lu = fmaf(a11, dx, fmaf(a12, dy, cu));
        const float lv = fmaf(a21, dx, fmaf(a22, dy, cv));
        const float fu = floorf(lu), fv = floorf(lv);
        float wu[4], wv[4], du[4], dv[4];
        keys_w(lu - fu, wu);
        keys_w(lv - fv, wv);
        keys_dw(lu - fu, du);
        keys_dw(lv - fv, dv);
        const int r0 = ou + (int)fu - 1, c0 = ov + (int)fv - 1;  // 1-based first tap
        float g = 0.f, gdu = 0.f, gdv = 0.f;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int y = min(max(c0 + q, 1), fr.W);
            const float *col = fr.g + (size_t)(y - 1) * fr.H;
            float sv = 0.f, sd = 0.f;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const float t = col[min(max(r0 + r, 1), fr.H) - 1];
                sv = fmaf(wu[r], t, sv);
                sd = fmaf(du[r], t, sd);
            }
            g = fmaf(wv[q], sv, g);
            gdu = fmaf(wv[q], sd, gdu);
            gdv = fmaf(dv[q], sv, gdv);
        }
        gv[p] = g;
        gu[p] = gdu;
        gw[p] = gdv;
        const double f = fidx(fr.f, fr.H, xa + i, ya + j);
        s[0] += f;
        s[1] += f * f;
        s[2] += g;
        s[3] += (double)g * g;
        s[4] += f * g;
    }
    double o[5];
    block_sum<BS, 5>(s, redd, o);
    if (tid == 0) {
        const double n = npix;
        const double mf = o[0] / n, mg = o[2] / n;
        const double bf = sqrt(fmax(o[1] - o[0] * o[0] / n, 0.0));
        const double bg = sqrt(fmax(o[3] - o[2] * o[2] / n, 0.0));
        const double zncc = (o[4] - o[0] * o[2] / n) / (bf * bg);
        stat[0] = mf;
        stat[1] = bf;
        stat[2] = mg;
        stat[3] = bg;
        stat[4] = zncc;
        if (a.cost) a.cost[idx] = 2.0 - 2.0 * zncc;
        if (a.status) a.status[idx] = 0;
    }
    __syncthreads();
    if (!a.gradP) return;
    const double mf = stat[0], bf = stat[1], mg = stat[2], bg = stat[3], zncc = stat[4];
    const double up = a.gcost ? a.gcost[idx] : 1.0;
    double gacc[6] = {0, 0, 0, 0, 0, 0};
    for (int p = tid; p < npix; p += BS) {
        const int j = (int)(((float)p + 0.5f) * rcp), i = p - j * nx;
        const double fh = ((double)fidx(fr.f, fr.H, xa + i, ya + j) - mf) / bf;
        const double gh = ((double)gv[p] - mg) / bg;
        const double dCdg = -(2.0 / bg) * (fh - zncc * gh);
        const double au = dCdg * gu[p], av = dCdg * gw[p];
        const double dx = i - (double)h, dy = j - (double)h;
        gacc[0] += au * dx;
        gacc[1] += av * dx;
        gacc[2] += au * dy;
        gacc[3] += av * dy;
        gacc[4] += au;
        gacc[5] += av;
    }
    double go[6];
    block_sum<BS, 6>(gacc, redd, go);
    if (tid == 0)
        for (int k = 0; k < 6; ++k) a.gradP[6 * idx + k] = up * go[k];
}

// The same cost and vector-Jacobian product with one wavefront per subset
// (windows of up to 64 * PPT pixels): the warped window's support of g is
// staged in an LDS tile with coalesced column loads, every lane gathers PPT
// pixels from it (g and, for the backward pass, its two derivatives from the
// same 16 taps), the five moments are FP32 per lane about c = f at the
// centre (small terms, so FP32 keeps them to a few ulps of the window's
// spread) and FP64 across the wavefront, and the gradient sums follow from
// the pixels the lanes still hold.  A subset whose support outgrows the tile
// gathers from HBM.
template <int PPT, bool GRAD>
__global__ __launch_bounds__(64) void znssd_wave_kernel(ZnssdArgs a, int tile_cap)
{
    extern __shared__ __align__(16) unsigned char smem[];
    float *tile = reinterpret_cast<float *>(smem);
    const long idx = blockIdx.x;
    const int lane = threadIdx.x;
    const FrameDev fr = load_frame(a.frames, a.slot ? gld(a.slot, (size_t)idx) : 0);
    const int w = gld(a.winsize, (size_t)idx), nx = w + 1, npix = nx * nx;
    const double x0 = round(gld(a.coords, (size_t)idx)), y0 = round(gld(a.coords, (size_t)(idx + a.N)));
    double P[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) P[k] = gld(a.P, (size_t)(6 * idx + k));
    const double h = w * 0.5;
    // same window rule as the IC-GN kernels (gradient array bounds), and the
    // warped corners inside the image test (funICGNQuadtree.m:149-157)
    bool ok = npix <= 64 * PPT && x0 - w / 2 >= 4 && y0 - w / 2 >= 4 && x0 + w / 2 <= fr.H - 3 &&
              y0 + w / 2 <= fr.W - 3;
    double umin = 1e300, umax = -1e300, vmin = 1e300, vmax = -1e300;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double dx = (q & 1) ? h : -h, dy = (q & 2) ? h : -h;
        const double u = (1 + P[0]) * dx + P[2] * dy + (x0 + P[4]);
        const double v = P[1] * dx + (1 + P[3]) * dy + (y0 + P[5]);
        ok = ok && u >= 3 && u <= fr.H - 2 && v >= 3 && v <= fr.W - 2;
        umin = fmin(umin, u);
        umax = fmax(umax, u);
        vmin = fmin(vmin, v);
        vmax = fmax(vmax, v);
    }
    if (!ok) {
        if (lane == 0) {
            if (a.cost) a.cost[idx] = 0.0;
            if (a.status) a.status[idx] = 1;
        }
        if (GRAD && a.gradP && lane < 6) a.gradP[6 * idx + lane] = 0.0;
        return;
    }
    const int xa = (int)x0 - w / 2, ya = (int)y0 - w / 2;
    // local warp relative to an integer origin near the window keeps the FP32
    // coordinates small
    const int ou = (int)floor(x0 + P[4]), ov = (int)floor(y0 + P[5]);
    const float a11 = (float)(1 + P[0]), a12 = (float)P[2], a21 = (float)P[1], a22 = (float)(1 + P[3]);
    const float cu = (float)((x0 + P[4]) - ou), cv = (float)((y0 + P[5]) - ov);
    // the taps' support: rows floor(u) - 1 .. floor(u) + 2 (one row of slack
    // for the FP32 local coordinates)
    const int tu0 = (int)floor(umin) - 2, tv0 = (int)floor(vmin) - 2;
    const int tnx = (int)floor(umax) + 4 - tu0, tny = (int)floor(vmax) + 4 - tv0;
    const bool use_tile = tnx * tny <= tile_cap;
    if (use_tile) {
        for (int t = lane; t < tnx * tny; t += 64) {
            const int sj = t / tnx, ri = t - sj * tnx;
            const int y = min(max(tv0 + sj, 1), fr.W), x = min(max(tu0 + ri, 1), fr.H);
            tile[t] = gld(fr.g, (size_t)((x - 1) + (y - 1) * fr.H));
        }
        __syncthreads();
    }
    const float rcp = 1.0f / nx, hf = w * 0.5f;
    const float c0 = fidx(fr.f, fr.H, xa + nx / 2, ya + nx / 2);
    float fv[PPT], gv[PPT], gu[PPT], gw[PPT];
    float s[5] = {0.f, 0.f, 0.f, 0.f, 0.f};  // about c0: f, f^2, g, g^2, f g
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        const int p = lane + 64 * k;
        fv[k] = gv[k] = gu[k] = gw[k] = 0.f;
        if (64 * k >= npix) continue;  // no lane holds a pixel of this slot
        if (p < npix) {
            const int j = (int)(((float)p + 0.5f) * rcp), i = p - j * nx;
            const float dx = (float)i - hf, dy = (float)j - hf;
            const float lu = fmaf(a11, dx, fmaf(a12, dy, cu));
            const float lv = fmaf(a21, dx, fmaf(a22, dy, cv));
            const float fu = floorf(lu), fvv = floorf(lv);
            float wu[4], wv[4], du[4], dv[4];
            keys_w(lu - fu, wu);
            keys_w(lv - fvv, wv);
            if (GRAD) {
                keys_dw(lu - fu, du);
                keys_dw(lv - fvv, dv);
            }
            const int r0 = ou + (int)fu - 1, cc0 = ov + (int)fvv - 1;  // 1-based first tap
            float g = 0.f, gdu = 0.f, gdv = 0.f;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                float t4[4];
                if (use_tile) {
                    const float *row = tile + (r0 - tu0) + (cc0 + q - tv0) * tnx;
#pragma unroll
                    for (int r = 0; r < 4; ++r) t4[r] = row[r];
                } else {
                    const int y = min(max(cc0 + q, 1), fr.W);
                    const float *col = fr.g + (size_t)(y - 1) * fr.H;
#pragma unroll
                    for (int r = 0; r < 4; ++r) t4[r] = gld(col, (size_t)(min(max(r0 + r, 1), fr.H) - 1));
                }
                float sv = 0.f, sd = 0.f;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    sv = fmaf(wu[r], t4[r], sv);
                    if (GRAD) sd = fmaf(du[r], t4[r], sd);
                }
                g = fmaf(wv[q], sv, g);
                if (GRAD) {
                    gdu = fmaf(wv[q], sd, gdu);
                    gdv = fmaf(dv[q], sv, gdv);
                }
            }
            const float f = fidx(fr.f, fr.H, xa + i, ya + j);
            fv[k] = f;
            gv[k] = g;
            gu[k] = gdu;
            gw[k] = gdv;
            const float ef = f - c0, eg = g - c0;
            s[0] += ef;
            s[1] = fmaf(ef, ef, s[1]);
            s[2] += eg;
            s[3] = fmaf(eg, eg, s[3]);
            s[4] = fmaf(ef, eg, s[4]);
        }
    }
    double m[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) m[q] = wave_sum_dpp_f64((double)s[q]);
    const double n = npix;
    const double df = m[0] / n, dg = m[2] / n;
    const double bf = sqrt(fmax(m[1] - m[0] * df, 0.0));
    const double bg = sqrt(fmax(m[3] - m[2] * dg, 0.0));
    const double zncc = (m[4] - m[0] * dg) / (bf * bg);
    if (!GRAD) {
        if (lane == 0) {
            if (a.cost) a.cost[idx] = 2.0 - 2.0 * zncc;
            if (a.status) a.status[idx] = 0;
        }
        return;
    }
    // dC/dg_p = -(2/bg) (fhat_p - ZNCC ghat_p)
    const float mf = (float)((double)c0 + df), mg = (float)((double)c0 + dg);
    const float ibf = (float)(1.0 / bf), ibg = (float)(1.0 / bg), z = (float)zncc;
    const float k2 = (float)(-2.0 / bg);
    float gs[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        const int p = lane + 64 * k;
        if (64 * k >= npix) continue;
        if (p < npix) {
            const int j = (int)(((float)p + 0.5f) * rcp), i = p - j * nx;
            const float dx = (float)i - hf, dy = (float)j - hf;
            const float fh = (fv[k] - mf) * ibf, gh = (gv[k] - mg) * ibg;
            const float dc = k2 * fmaf(-z, gh, fh);
            const float au = dc * gu[k], av = dc * gw[k];
            gs[0] = fmaf(au, dx, gs[0]);
            gs[1] = fmaf(av, dx, gs[1]);
            gs[2] = fmaf(au, dy, gs[2]);
            gs[3] = fmaf(av, dy, gs[3]);
            gs[4] += au;
            gs[5] += av;
        }
    }
    const double up = a.gcost ? gld(a.gcost, (size_t)idx) : 1.0;
    double go[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) go[q] = wave_sum_dpp_f64((double)gs[q]);
    if (lane == 0) {
        for (int q = 0; q < 6; ++q) a.gradP[6 * idx + q] = up * go[q];
        if (a.cost) a.cost[idx] = 2.0 - 2.0 * zncc;
        if (a.status) a.status[idx] = 0;
    }
}

// Zeros of f * mask in each subset's window (LocalICGNQuadtree.m:71-82): a
// window with any grows at set-up (funICGNQuadtree.m:66-70), so the batch
// bins such subsets by their grown size, apart from the rest.  One wavefront
// per subset; -1 for a window that does not fit in the image.
__global__ __launch_bounds__(64) void mask_class_kernel(const FrameDev *frames, const int *slot,
                                                         const double *coords, const int *wx,
                                                         const int *wy, long N, int *out)
{
    const long i = blockIdx.x;
    if (i >= N) return;
    const FrameDev fr = frames[slot ? slot[i] : 0];
    const int w = wx[i], h = wy[i];
    const int xa = (int)round(coords[i]) - w / 2, ya = (int)round(coords[i + N]) - h / 2;
    if (xa < 1 || ya < 1 || xa + w > fr.H || ya + h > fr.W) {
        if (threadIdx.x == 0) out[i] = -1;
        return;
    }
    const int nx = w + 1, npix = nx * (h + 1);
    int z = 0;
    for (int p = threadIdx.x; p < npix; p += 64) {
        const int j = p / nx, q = p - j * nx;
        z += fidx(fr.f, fr.H, xa + q, ya + j) * fidx(fr.mask, fr.H, xa + q, ya + j) == 0.f;
    }
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) z += __shfl_xor(z, o, 64);
    if (threadIdx.x == 0) out[i] = z;
}

hipError_t mask_class_launch(const FrameDev *frames, const int *slot, const double *coords,
                             const int *wx, const int *wy, long N, int *out, hipStream_t s)
{
    if (N <= 0) return hipSuccess;
    hipLaunchKernelGGL(mask_class_kernel, dim3((unsigned)N), dim3(64), 0, s, frames, slot, coords,
                       wx, wy, N, out);
    return hipGetLastError();
}

// windows of up to 32 px: one wavefront per subset; larger: one workgroup
template <int PPT>
static hipError_t znssd_wave(const ZnssdArgs &a, int wmax, hipStream_t s)
{
    // the warped support of the largest window with a 12 % stretch, plus
    // the taps and their slack
    const int side = (int)ceil((wmax + 1) * 1.12) + 8;
    const int tile_cap = side * side;
    const size_t lds = sizeof(float) * (size_t)tile_cap;
    if (a.gradP)
        hipLaunchKernelGGL((znssd_wave_kernel<PPT, true>), dim3((unsigned)a.N), dim3(64), lds, s,
                           a, tile_cap);
    else
        hipLaunchKernelGGL((znssd_wave_kernel<PPT, false>), dim3((unsigned)a.N), dim3(64), lds, s,
                           a, tile_cap);
    return hipGetLastError();
}

hipError_t znssd_launch(const ZnssdArgs &a, hipStream_t s)
{
    if (a.N <= 0) return hipSuccess;
    const int npix = a.npix_cap;
    if (npix <= 64 * 5) return znssd_wave<5>(a, (int)std::sqrt((double)npix) - 1, s);
    if (npix <= 64 * 9) return znssd_wave<9>(a, (int)std::sqrt((double)npix) - 1, s);
    if (npix <= 64 * 18) return znssd_wave<18>(a, (int)std::sqrt((double)npix) - 1, s);
    const size_t lds = 3 * sizeof(float) * (size_t)a.npix_cap;
    static bool attr = false;
    if (!attr) {
        hipFuncSetAttribute((const void *)znssd_kernel<256>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 8192);
        attr = true;
    }
    hipLaunchKernelGGL(znssd_kernel<256>, dim3((unsigned)a.N), dim3(256), lds, s, a);
    return hipGetLastError();
}

}  // namespace staq
