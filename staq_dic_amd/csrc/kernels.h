// Kernel argument blocks and launchers shared by the kernels and the C ABI.
#pragma once

#include "common.h"

namespace staq {

struct IcgnArgs {
    const FrameDev *frames;
    const int *slot;        // frame index per subset (nullptr -> 0)
    const double *coords;   // N x 2 column-major (coordinatesFEM)
    const int *wx, *wy;     // window size per subset (wy: 2-DOF only)
    const int *order;       // processing order (nullptr -> identity)
    const double *U0;       // 6-DOF: initial displacement, 2N interleaved
    const double *UOld;     // 2-DOF: U of the global step, 2N
    const double *FOld;     // 2-DOF: F of the global step, 4N
    const double *vdual;    // 2-DOF: displacement dual, 2N
    double *U, *F, *H21;    // outputs (F, H21: 6-DOF)
    int *status;
    int *work_counter;      // persistent mode: device queue head (nullptr -> 1 subset/block)
    const int *defer_flag;  // fallback launch: number of subsets listed for it
    int *defer_out;         // main launch: count of subsets listed for the small fallback launch
    int *defer_out2;        // main launch: count of subsets listed for the large fallback launch
    int *defer_list;        // N entries: subsets for the small fallback launch from the front
                            // (windows of up to kSmallFallbackW px), the others from the back
    int list_back;          // fallback launch: take the list from the back
    long N;                 // subset count (array extents)
    long n_items;           // items of this launch
    long item0;             // first item (one launch per subset uses this)
    double tol, mu;
    int plain;              // uniform-grid semantics (no node rounding, whole-window stats)
    int max_it;             // 100 in the reference
    int npix_cap, tile_cap; // dynamic LDS sizing (pixels, floats)
    int defer;              // a window larger than npix_cap is left for a fallback launch
    int only_deferred;      // fallback launch: process only subsets left by the main launch
    int *handed;            // lean kernel: set to 1 for a subset it leaves to the fallback (may be null)
    double guard2;          // guard band of the relative stopping test, on squared norms (0 = none)
    double guard2_abs;      // guard band of the absolute stopping test |b| < tol (0 = none)
    int *exact_count;       // incremented for each subset left to the FP64 re-solve (may be null)
    int grad_lds;           // FP64 re-solve: stage the window's DfDx, DfDy in LDS once per subset
    // the guard bands as limits on the squared norms, tol^2 (1 -+ guard2) and
    // tol^2 (1 -+ guard2_abs) (lean kernel; +inf / -inf when the guard is off)
    double rel_lo2, rel_hi2, abs_lo2, abs_hi2;
};

struct ZnssdArgs {
    const FrameDev *frames;
    const int *slot;
    const double *coords;  // N x 2 column-major
    const int *winsize;
    const double *P;       // 6N: [u_x v_x u_y v_y u v] per subset
    const double *gcost;   // upstream dL/dC_i (backward), may be null (= 1)
    double *cost;          // N (forward), may be null
    double *gradP;         // 6N (backward), may be null
    int *status;           // N: 0 ok, 1 window or warp outside the image
    long N;
    int npix_cap;
};

struct SearchArgs {
    const float *f, *g;
    int H, W;
    int r1, r2, w;
    int gx0, gy0, step, nx, ny;  // grid: ii = gx0 + ix*step, jj = gy0 + iy*step
    double2 *scratch;            // (num, den) per kept shift, per point of this launch
    double2 *colsum;             // strip kernel: column box sums, k1 x (w+2r2+3) per point
    long point0;                 // first point of this launch
    double *x0, *y0, *u, *v, *ccmax;  // ny x nx column-major outputs
};

// Windows of up to kSmallFallbackW px on each side grow to at most 47 x 47 px
// (funICGNQuadtree.m:68-79, funICGN_Subpb1.m:78-79, funICGN.m:58), so the
// fallback launch for them needs room for 48 x 48 pixels only.
constexpr int kSmallFallbackW = 32;
constexpr int kSmallFallbackPix = 48 * 48;

// A main launch leaves subset idx (given window wx x wy) to a fallback launch.
__device__ __forceinline__ void defer_append(const IcgnArgs &a, long idx, int wx, int wy)
{
    if (wx <= kSmallFallbackW && wy <= kSmallFallbackW) {
        if (!a.defer_out) return;
        const int pos = atomicAdd(a.defer_out, 1);
        if (a.defer_list) a.defer_list[pos] = (int)idx;
    } else {
        if (!a.defer_out2) return;
        const int pos = atomicAdd(a.defer_out2, 1);
        if (a.defer_list) a.defer_list[a.N - 1 - pos] = (int)idx;
    }
}

hipError_t icgn_launch(const IcgnArgs &a, int mode, int bs, int grid, hipStream_t s);
// wave-per-subset kernel: pixels per lane for a window of npix_cap pixels (0 = too large)
// pixels per lane for windows of npix_cap pixels spread over nw wavefronts
// (0: too large for the wavefront kernel)
int icgn_wave_ppt(int npix_cap, int nw);
hipError_t icgn_wave_launch(const IcgnArgs &a, int mode, int ppt, int nw, int grid, hipStream_t s);
// lean wavefront kernel for windows without masked pixels (icgn_lean.hip),
// with 1, 2, 4, 8 or 16 wavefronts per subset; pixels per lane as
// icgn_wave_ppt up to four wavefronts (0: too large)
int icgn_lean_ppt(int npix_cap, int nw);
hipError_t icgn_lean_launch(const IcgnArgs &a, int mode, int ppt, int nw, int grid, hipStream_t s);
size_t icgn_lds_bytes(int npix_cap, int tile_cap, bool exact = false, bool grad = false);
// FP64 re-solve (workgroup kernel, 256 threads) of the subsets a launch left
// with status ST_EXACT or ST_DEFER (only_deferred set); g from an LDS tile
// when tile_cap > 0, else from HBM
hipError_t icgn_exact_launch(const IcgnArgs &a, int mode, int bs, int grid, hipStream_t s);
hipError_t ba_interp2_launch(const void *F, long M, long N, long P, const void *X, const void *Y,
                             long nd, void *out, bool single, hipStream_t s);
hipError_t img_gradient_launch(const float *f, const float *mask, int H, int W, float *dfdx,
                               float *dfdy, hipStream_t s);
hipError_t znssd_launch(const ZnssdArgs &a, hipStream_t s);
hipError_t mask_class_launch(const FrameDev *frames, const int *slot, const double *coords,
                             const int *wx, const int *wy, long N, int *out, hipStream_t s);
hipError_t search_launch(const SearchArgs &a, long npoints, hipStream_t s);
size_t search_lds_bytes(int w, int r1, int r2);
// the register-blocked strip kernel takes geometries whose padded blocks fit its LDS budget
size_t search_strip_lds_bytes(int w, int r1, int r2);
bool search_uses_strip(int w, int r1, int r2);

// multigrid search (mgsearch.hip).  Level 1: the full normxcorr2 map of
// f(box_f) against g(box_g) and findpeak on it; four device buffers of the
// sizes xcorr_full_sizes gives; out = xpeak (dim 2), ypeak (dim 1), max_f
// (NaN for a constant template).  Returns 0 or a negative error with text.
void xcorr_full_sizes(long n1, long n2, long a1, long a2, size_t bytes[4]);
int xcorr_full(const FrameDev &fr, const long box_f[4], const long box_g[4], void *const buf[4],
               hipStream_t s, double out[3], char *err, size_t errlen);
// Levels >= 2: nb boxes (device arrays: bf, bg 4 per box; tile0 prefix sums of
// xcorr_box_tiles), part = ntile * xcorr_box_part_doubles(pad) doubles
size_t xcorr_box_part_doubles(int pad);
// Quality factors (pce, ppe) of the whole normxcorr2 map of each box
// (qfactor.hip): boxes_lds run with C and D in lds_bytes of LDS, boxes_hbm
// read them from the frame; map holds 2 doubles per placement from map0[b].
size_t qfactor_lds_bytes(long n1, long n2, long d1, long d2);
hipError_t qfactor_launch(const FrameDev &fr, const long *bf, const long *bg, const long *map0,
                          double *map, double *q, int *ok, const long *boxes_lds, long n_lds,
                          size_t lds_bytes, const long *boxes_hbm, long n_hbm, hipStream_t s);
int xcorr_box_tiles(long n1, long n2);
hipError_t xcorr_boxes_launch(const FrameDev &fr, int pad, long nb, long ntile, const long *bf,
                              const long *bg, const int *tile0, double *part, double *out, int *ok,
                              hipStream_t s);

}  // namespace staq
