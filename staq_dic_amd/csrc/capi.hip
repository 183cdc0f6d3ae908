// Host side of the C ABI declared in include/staq_dic.h.
//
// A context owns one HIP stream (default priority) and the frames resident in
// HBM.  A batch owns a device-resident subset list sorted into window-size
// bins; a local step is: reset kernel (queue heads, defer flag), one launch
// per bin with LDS sized for that bin, and a fallback launch that only runs
// if some subset's window grew beyond its bin (masked subsets,
// funICGNQuadtree.m:68-79).  Everything a step enqueues is a kernel, so the
// step can be captured into a hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/staq_dic.h"
#include "kernels.h"



using namespace staq;

// ---------------------------------------------------------------------------
// small device helpers
// ---------------------------------------------------------------------------
__global__ void reset_ints_kernel(int *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Device memory of staq_local_icgn_noalloc, one copy per GPU for the life of
// the process: the step's counters (queue heads and list lengths of the
// fallback launches, the count of FP64 re-solves) in the first 64 ints, then
// the defer list, one int per subset. The name is a host-side stand-in: host
// code uses only the address hipGetSymbolAddress returns for it on each
// device (staq_ctx::noalloc).
constexpr long kNoAllocMaxN = 1L << 20;  // subsets per call
constexpr int kNoAllocHead = 64;         // ints before the defer list
static __device__ __align__(256) int scratch[kNoAllocHead + kNoAllocMaxN];

__global__ void chain_kernel(const double *U, const double *F, double *UOld, double *FOld,
                             double *vdual, long N)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < 2 * N) {
        UOld[i] = U[i];
        vdual[i] = 0.0;
    }
    if (i < 4 * N) FOld[i] = F[i];
}

__global__ void f64_to_f32_kernel(const double *in, float *out, long n)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)in[i];
}

__global__ void ones_kernel(float *p, long n)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) p[i] = 1.f;
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
namespace {

struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
    hipError_t ensure(size_t bytes)
    {
        if (bytes <= n) return hipSuccess;
        if (p) hipFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
        if (e == hipSuccess) n = bytes;
        return e;
    }
    void release()
    {
        if (p) hipFree(p);
        p = nullptr;
        n = 0;
    }
    template <typename T>
    T *as() const { return reinterpret_cast<T *>(p); }
};

struct Frame {
    DevBuf f, g, mask, dfdx, dfdy;
    long H = 0, W = 0;
    bool set = false;
};

}  // namespace

struct staq_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::vector<Frame> frames;
    DevBuf d_frames;
    DevBuf s[12];  // scratch for host-pointer calls
    int *noalloc = nullptr;  // this device's copy of `scratch` (staq_local_icgn_noalloc)
};

struct Bin {
    int wxcap = 0, wycap = 0;
    long item0 = 0, n = 0;
    int npix_cap = 0, tile_cap = 0;
    // [mode 6, mode 2] kernel form: 0 = one wavefront per subset, -n = n
    // wavefronts per subset, > 0 = workgroup size of the workgroup kernel
    int bs[2] = {0, 0};
    int ppt = 0;         // pixels per lane of the one-wavefront kernel (0: not eligible)
    int wave_npix = 0;   // window pixels the wavefront kernels must hold (grown windows included)
    int masked = 0;      // the bin's windows touch the mask (their window grows)
};

struct staq_batch {
    staq_ctx *ctx = nullptr;
    long N = 0;
    DevBuf coords, U0, UOld, FOld, vdual, U, F, H, status, wx, wy, slot, order, ints, handed;
    DevBuf deflist;  // subsets left to the fallback launch in the current step
    std::vector<Bin> bins;
    // host copies behind the bins: window sizes, zeros of f*mask in each
    // window, the grown window's pixels and pixel slots, and whether the
    // subset goes to the general kernel (mask zeros, or handed back by the
    // lean kernel in an earlier step)
    std::vector<int> hx, hy, cnt, gnp, gppt;
    std::vector<char> general;
    int wmax = 0;
    int fb_npix_cap = 0, fb_tile_cap = 0;
    bool has_slot = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int graph_mode = 0;
    // bins run concurrently: one side stream per extra bin, forked from and
    // joined into the context stream with events
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> ev;  // [0] fork, [1..] joins
    // timing events (created on first timed step): a start/stop pair per bin
    // launch and one for the fallback launch, each recorded on that launch's stream
    std::vector<hipEvent_t> tev;
};

static int fail(staq_ctx *ctx, int code, const std::string &msg)
{
    if (ctx) ctx->err = msg;
    return code;
}

static int hipfail(staq_ctx *ctx, hipError_t e, const char *what)
{
    return fail(ctx, e == hipErrorOutOfMemory ? STAQ_ENOMEM : STAQ_EHIP,
                std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCHK(ctx, x)                                  \
    do {                                                \
        hipError_t e_ = (x);                            \
        if (e_ != hipSuccess) return hipfail(ctx, e_, #x); \
    } while (0)

static hipStream_t pick_stream(staq_ctx *ctx, void *s)
{
    return s ? reinterpret_cast<hipStream_t>(s) : ctx->stream;
}

extern "C" {

void staq_params_default(staq_params *p)
{
    if (!p) return;
    p->tol = 1e-2;
    p->mu = 1e-3;
    p->max_iter = 100;
    p->plain = 0;
    p->block_size = 0;
    p->persistent = 0;
    p->guard = 2e-3;
}

int staq_ctx_create(staq_ctx **out, int device)
{
    if (!out) return STAQ_EINVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return STAQ_EHIP;
    if (device < 0 || device >= n) return STAQ_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return STAQ_EHIP;
    staq_ctx *c = new staq_ctx();
    c->device = device;
    if (hipStreamCreate(&c->stream) != hipSuccess) {
        delete c;
        return STAQ_EHIP;
    }
    // this device's copy of the no-alloc entry's memory (device current)
    if (hipGetSymbolAddress((void **)&c->noalloc, HIP_SYMBOL(scratch)) != hipSuccess ||
        !c->noalloc) {
        hipStreamDestroy(c->stream);
        delete c;
        return STAQ_EHIP;
    }
    *out = c;
    return STAQ_OK;
}

void staq_ctx_destroy(staq_ctx *ctx)
{
    if (!ctx) return;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    for (auto &f : ctx->frames) {
        f.f.release();
        f.g.release();
        f.mask.release();
        f.dfdx.release();
        f.dfdy.release();
    }
    ctx->d_frames.release();
    for (auto &b : ctx->s) b.release();
    hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *staq_last_error(const staq_ctx *ctx) { return ctx ? ctx->err.c_str() : "no context"; }

void *staq_ctx_stream(staq_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

}  // extern "C"

// Refresh the device array of frame descriptors.
static int sync_frames(staq_ctx *ctx, hipStream_t s)
{
    std::vector<FrameDev> h(ctx->frames.size());
    for (size_t i = 0; i < h.size(); ++i) {
        const Frame &f = ctx->frames[i];
        h[i].f = f.f.as<float>();
        h[i].g = f.g.as<float>();
        h[i].mask = f.mask.as<float>();
        h[i].dfdx = f.dfdx.as<float>();
        h[i].dfdy = f.dfdy.as<float>();
        h[i].H = (int)f.H;
        h[i].W = (int)f.W;
    }
    HIPCHK(ctx, ctx->d_frames.ensure(sizeof(FrameDev) * std::max<size_t>(h.size(), 1)));
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_frames.p, h.data(), sizeof(FrameDev) * h.size(),
                               hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    return STAQ_OK;
}

static int frame_alloc(staq_ctx *ctx, int slot, long H, long W)
{
    if (slot < 0 || slot > 4096) return fail(ctx, STAQ_EINVAL, "frame slot out of range");
    if (H < 16 || W < 16 || H > (1L << 20) || W > (1L << 20))
        return fail(ctx, STAQ_EINVAL, "image must be at least 16 x 16");
    if ((size_t)slot >= ctx->frames.size()) ctx->frames.resize(slot + 1);
    Frame &fr = ctx->frames[slot];
    const size_t n = (size_t)H * W, nd = (size_t)(H - 6) * (W - 6);
    HIPCHK(ctx, fr.f.ensure(n * 4));
    HIPCHK(ctx, fr.g.ensure(n * 4));
    HIPCHK(ctx, fr.mask.ensure(n * 4));
    HIPCHK(ctx, fr.dfdx.ensure(nd * 4));
    HIPCHK(ctx, fr.dfdy.ensure(nd * 4));
    fr.H = H;
    fr.W = W;
    return STAQ_OK;
}

static int frame_finish(staq_ctx *ctx, int slot, hipStream_t s)
{
    Frame &fr = ctx->frames[slot];
    HIPCHK(ctx, img_gradient_launch(fr.f.as<float>(), fr.mask.as<float>(), (int)fr.H, (int)fr.W,
                                    fr.dfdx.as<float>(), fr.dfdy.as<float>(), s));
    fr.set = true;
    return sync_frames(ctx, s);
}

extern "C" {

int staq_set_frame_f32(staq_ctx *ctx, int slot, const float *f, const float *g, const float *mask,
                       long H, long W)
{
    if (!ctx || !f || !g) return fail(ctx, STAQ_EINVAL, "null argument");
    hipSetDevice(ctx->device);
    int rc = frame_alloc(ctx, slot, H, W);
    if (rc) return rc;
    Frame &fr = ctx->frames[slot];
    const size_t n = (size_t)H * W;
    HIPCHK(ctx, hipMemcpyAsync(fr.f.p, f, n * 4, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(fr.g.p, g, n * 4, hipMemcpyHostToDevice, ctx->stream));
    if (mask)
        HIPCHK(ctx, hipMemcpyAsync(fr.mask.p, mask, n * 4, hipMemcpyHostToDevice, ctx->stream));
    else
        hipLaunchKernelGGL(ones_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream,
                           fr.mask.as<float>(), (long)n);
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return frame_finish(ctx, slot, ctx->stream);
}

int staq_set_frame(staq_ctx *ctx, int slot, const double *f, const double *g, const double *mask,
                   long H, long W)
{
    if (!ctx || !f || !g) return fail(ctx, STAQ_EINVAL, "null argument");
    if (H < 16 || W < 16) return fail(ctx, STAQ_EINVAL, "image must be at least 16 x 16");
    const size_t n = (size_t)H * W;
    std::vector<float> ff(n), gg(n), mm;
    for (size_t i = 0; i < n; ++i) {
        ff[i] = (float)f[i];
        gg[i] = (float)g[i];
    }
    if (mask) {
        mm.resize(n);
        for (size_t i = 0; i < n; ++i) mm[i] = (float)mask[i];
    }
    return staq_set_frame_f32(ctx, slot, ff.data(), gg.data(), mask ? mm.data() : nullptr, H, W);
}

int staq_set_frame_dev(staq_ctx *ctx, int slot, const float *f, const float *g, const float *mask,
                       long H, long W, void *stream)
{
    if (!ctx || !f || !g) return fail(ctx, STAQ_EINVAL, "null argument");
    hipSetDevice(ctx->device);
    hipStream_t s = pick_stream(ctx, stream);
    int rc = frame_alloc(ctx, slot, H, W);
    if (rc) return rc;
    Frame &fr = ctx->frames[slot];
    const size_t n = (size_t)H * W;
    HIPCHK(ctx, hipMemcpyAsync(fr.f.p, f, n * 4, hipMemcpyDeviceToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(fr.g.p, g, n * 4, hipMemcpyDeviceToDevice, s));
    if (mask)
        HIPCHK(ctx, hipMemcpyAsync(fr.mask.p, mask, n * 4, hipMemcpyDeviceToDevice, s));
    else
        hipLaunchKernelGGL(ones_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                           fr.mask.as<float>(), (long)n);
    return frame_finish(ctx, slot, s);
}

int staq_get_gradient(staq_ctx *ctx, int slot, double *dfdx, double *dfdy)
{
    if (!ctx || slot < 0 || (size_t)slot >= ctx->frames.size() || !ctx->frames[slot].set)
        return fail(ctx, STAQ_ENOFRAME, "frame slot not set");
    hipSetDevice(ctx->device);
    Frame &fr = ctx->frames[slot];
    const size_t nd = (size_t)(fr.H - 6) * (fr.W - 6);
    std::vector<float> t(nd);
    HIPCHK(ctx, hipMemcpy(t.data(), fr.dfdx.p, nd * 4, hipMemcpyDeviceToHost));
    if (dfdx)
        for (size_t i = 0; i < nd; ++i) dfdx[i] = t[i];
    HIPCHK(ctx, hipMemcpy(t.data(), fr.dfdy.p, nd * 4, hipMemcpyDeviceToHost));
    if (dfdy)
        for (size_t i = 0; i < nd; ++i) dfdy[i] = t[i];
    return STAQ_OK;
}

// ---------------------------------------------------------------------------
// ba_interp2 / gradient drop-ins
// ---------------------------------------------------------------------------
int staq_ba_interp2_dev(staq_ctx *ctx, const void *F, long M, long N, long P, const void *X,
                        const void *Y, long nd, void *Z, int single, void *stream)
{
    if (!ctx || !F || !X || !Y || !Z || M <= 0 || N <= 0 || P <= 0 || nd < 0)
        return fail(ctx, STAQ_EINVAL, "bad ba_interp2 arguments");
    hipSetDevice(ctx->device);
    HIPCHK(ctx, ba_interp2_launch(F, M, N, P, X, Y, nd, Z, single != 0, pick_stream(ctx, stream)));
    return STAQ_OK;
}

static int interp_host(staq_ctx *ctx, const void *F, long M, long N, long P, const void *X,
                       const void *Y, long nd, void *Z, size_t es)
{
    if (!ctx || !F || !X || !Y || !Z || M <= 0 || N <= 0 || P <= 0 || nd < 0)
        return fail(ctx, STAQ_EINVAL, "bad ba_interp2 arguments");
    hipSetDevice(ctx->device);
    const size_t nF = (size_t)M * N * P * es, nX = (size_t)nd * es, nZ = (size_t)nd * P * es;
    HIPCHK(ctx, ctx->s[0].ensure(nF));
    HIPCHK(ctx, ctx->s[1].ensure(nX));
    HIPCHK(ctx, ctx->s[2].ensure(nX));
    HIPCHK(ctx, ctx->s[3].ensure(nZ));
    hipStream_t s = ctx->stream;
    HIPCHK(ctx, hipMemcpyAsync(ctx->s[0].p, F, nF, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ctx->s[1].p, X, nX, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ctx->s[2].p, Y, nX, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, ba_interp2_launch(ctx->s[0].p, M, N, P, ctx->s[1].p, ctx->s[2].p, nd, ctx->s[3].p,
                                  es == 4, s));
    HIPCHK(ctx, hipMemcpyAsync(Z, ctx->s[3].p, nZ, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    return STAQ_OK;
}

int staq_ba_interp2(staq_ctx *ctx, const double *F, long M, long N, long P, const double *X,
                    const double *Y, long nd, double *Z)
{
    return interp_host(ctx, F, M, N, P, X, Y, nd, Z, 8);
}

int staq_ba_interp2_f32(staq_ctx *ctx, const float *F, long M, long N, long P, const float *X,
                        const float *Y, long nd, float *Z)
{
    return interp_host(ctx, F, M, N, P, X, Y, nd, Z, 4);
}

int staq_img_gradient(staq_ctx *ctx, const double *img, const double *mask, long H, long W,
                      double *dfdx, double *dfdy)
{
    if (!ctx || !img || !dfdx || !dfdy || H < 7 || W < 7)
        return fail(ctx, STAQ_EINVAL, "bad gradient arguments");
    hipSetDevice(ctx->device);
    const size_t n = (size_t)H * W, nd = (size_t)(H - 6) * (W - 6);
    std::vector<float> f(n), m(n, 1.f), ox(nd), oy(nd);
    for (size_t i = 0; i < n; ++i) f[i] = (float)img[i];
    if (mask)
        for (size_t i = 0; i < n; ++i) m[i] = (float)mask[i];
    HIPCHK(ctx, ctx->s[0].ensure(n * 4));
    HIPCHK(ctx, ctx->s[1].ensure(n * 4));
    HIPCHK(ctx, ctx->s[2].ensure(nd * 4));
    HIPCHK(ctx, ctx->s[3].ensure(nd * 4));
    hipStream_t s = ctx->stream;
    HIPCHK(ctx, hipMemcpyAsync(ctx->s[0].p, f.data(), n * 4, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ctx->s[1].p, m.data(), n * 4, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, img_gradient_launch(ctx->s[0].as<float>(), ctx->s[1].as<float>(), (int)H, (int)W,
                                    ctx->s[2].as<float>(), ctx->s[3].as<float>(), s));
    HIPCHK(ctx, hipMemcpyAsync(ox.data(), ctx->s[2].p, nd * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipMemcpyAsync(oy.data(), ctx->s[3].p, nd * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    for (size_t i = 0; i < nd; ++i) {
        dfdx[i] = ox[i];
        dfdy[i] = oy[i];
    }
    return STAQ_OK;
}

// ---------------------------------------------------------------------------
// integer search
// ---------------------------------------------------------------------------
int staq_integer_search(staq_ctx *ctx, int slot, long r1, long r2, const long gridx[2],
                        const long gridy[2], long w, long step, long max_pts, long *nx_out,
                        long *ny_out, double *x0, double *y0, double *u, double *v,
                        double *ccmax)
{
    if (!ctx || !gridx || !gridy || !nx_out || !ny_out || !x0 || !y0 || !u || !v || !ccmax)
        return fail(ctx, STAQ_EINVAL, "null argument");
    if (slot < 0 || (size_t)slot >= ctx->frames.size() || !ctx->frames[slot].set)
        return fail(ctx, STAQ_ENOFRAME, "frame slot not set");
    if (r1 < 0 || r2 < 0 || w < 2 || step < 1) return fail(ctx, STAQ_EINVAL, "bad search geometry");
    if (!search_uses_strip((int)w, (int)r1, (int)r2) &&
        search_lds_bytes((int)w, (int)r1, (int)r2) > 150 * 1024)
        return fail(ctx, STAQ_EINVAL, "template plus search block exceed LDS");
    hipSetDevice(ctx->device);
    const Frame &fr = ctx->frames[slot];
    const long H = fr.H, W = fr.W;
    // funIntegerSearch.m:14-25 ROI shrink, :90-91 grid
    long gx0 = gridx[0], gx1 = gridx[1], gy0 = gridy[0], gy1 = gridy[1];
    while (gx0 - r1 - 0.5 * w < 1) ++gx0;
    while (gy0 - r2 - 0.5 * w < 1) ++gy0;
    while (gx1 + 1.5 * w + r1 > H) --gx1;
    while (gy1 + 1.5 * w + r2 > W) --gy1;
    const long nx = (gx1 - step >= gx0) ? (gx1 - step - gx0) / step + 1 : 0;
    const long ny = (gy1 - step >= gy0) ? (gy1 - step - gy0) / step + 1 : 0;
    *nx_out = nx;
    *ny_out = ny;
    if (nx * ny > max_pts) return fail(ctx, STAQ_EINVAL, "output arrays too small for the grid");
    const long np = nx * ny;
    if (np == 0) return STAQ_OK;
    const long kk = (2 * r1 + 3) * (2 * r2 + 3);
    // per point: the (num, den) map and, for the strip kernel, the column box sums
    const long kcs = search_uses_strip((int)w, (int)r1, (int)r2)
                         ? (2 * r1 + 3) * (w + 2 * r2 + 3) : 0;
    const long chunk = std::max(1L, std::min(np, (512L << 20) / ((kk + kcs) * 16)));
    hipStream_t s = ctx->stream;
    HIPCHK(ctx, ctx->s[0].ensure((size_t)chunk * kk * 16));
    HIPCHK(ctx, ctx->s[6].ensure((size_t)chunk * std::max(kcs, 1L) * 16));
    for (int k = 0; k < 5; ++k) HIPCHK(ctx, ctx->s[1 + k].ensure((size_t)np * 8));
    SearchArgs a;
    a.f = fr.f.as<float>();
    a.g = fr.g.as<float>();
    a.H = (int)H;
    a.W = (int)W;
    a.r1 = (int)r1;
    a.r2 = (int)r2;
    a.w = (int)w;
    a.gx0 = (int)gx0;
    a.gy0 = (int)gy0;
    a.step = (int)step;
    a.nx = (int)nx;
    a.ny = (int)ny;
    a.scratch = ctx->s[0].as<double2>();
    a.colsum = ctx->s[6].as<double2>();
    a.x0 = ctx->s[1].as<double>();
    a.y0 = ctx->s[2].as<double>();
    a.u = ctx->s[3].as<double>();
    a.v = ctx->s[4].as<double>();
    a.ccmax = ctx->s[5].as<double>();
    for (long p0 = 0; p0 < np; p0 += chunk) {
        a.point0 = p0;
        HIPCHK(ctx, search_launch(a, std::min(chunk, np - p0), s));
    }
    double *outs[5] = {x0, y0, u, v, ccmax};
    for (int k = 0; k < 5; ++k)
        HIPCHK(ctx, hipMemcpyAsync(outs[k], ctx->s[1 + k].p, (size_t)np * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    return STAQ_OK;
}

static bool frame_desc(staq_ctx *ctx, int slot, FrameDev &d)
{
    if (slot < 0 || (size_t)slot >= ctx->frames.size() || !ctx->frames[slot].set) return false;
    const Frame &f = ctx->frames[slot];
    d.f = f.f.as<float>();
    d.g = f.g.as<float>();
    d.mask = f.mask.as<float>();
    d.dfdx = f.dfdx.as<float>();
    d.dfdy = f.dfdy.as<float>();
    d.H = (int)f.H;
    d.W = (int)f.W;
    return true;
}

int staq_xcorr_full_peak(staq_ctx *ctx, int slot, const long box_f[4], const long box_g[4],
                         double out[3])
{
    if (!ctx || !box_f || !box_g || !out) return fail(ctx, STAQ_EINVAL, "null argument");
    FrameDev fr;
    if (!frame_desc(ctx, slot, fr)) return fail(ctx, STAQ_ENOFRAME, "frame slot not set");
    const long n1 = box_f[1] - box_f[0] + 1, n2 = box_f[3] - box_f[2] + 1;
    const long a1 = box_g[1] - box_g[0] + 1, a2 = box_g[3] - box_g[2] + 1;
    if (n1 < 2 || n2 < 2 || a1 < n1 || a2 < n2 || box_f[0] < 1 || box_f[2] < 1 ||
        box_f[1] > fr.H || box_f[3] > fr.W || box_g[0] < 1 || box_g[2] < 1 || box_g[1] > fr.H ||
        box_g[3] > fr.W)
        return fail(ctx, STAQ_EINVAL, "box outside the image, or template larger than the image");
    hipSetDevice(ctx->device);
    size_t bytes[4];
    xcorr_full_sizes(n1, n2, a1, a2, bytes);
    void *buf[4];
    for (int k = 0; k < 4; ++k) {
        HIPCHK(ctx, ctx->s[6 + k].ensure(bytes[k]));
        buf[k] = ctx->s[6 + k].p;
    }
    char err[256] = {0};
    const int rc = xcorr_full(fr, box_f, box_g, buf, ctx->stream, out, err, sizeof err);
    if (rc) return fail(ctx, STAQ_EHIP, err);
    return STAQ_OK;
}

int staq_xcorr_boxes(staq_ctx *ctx, int slot, long nbox, const long *box_f, const long *box_g,
                     long pad, double *out, int *ok)
{
    if (!ctx || nbox < 0 || (nbox > 0 && (!box_f || !box_g || !out || !ok)))
        return fail(ctx, STAQ_EINVAL, "null argument");
    if (pad < 0 || pad > 6) return fail(ctx, STAQ_EINVAL, "pad must be 0..6");
    FrameDev fr;
    if (!frame_desc(ctx, slot, fr)) return fail(ctx, STAQ_ENOFRAME, "frame slot not set");
    if (nbox == 0) return STAQ_OK;
    hipSetDevice(ctx->device);
    // boxes whose g support leaves the image, or whose template and g box
    // differ in size, are the reference's catch (NaN); the others are computed
    std::vector<long> bf, bg, idx;
    std::vector<int> tile0(1, 0);
    for (long b = 0; b < nbox; ++b) {
        const long *F = box_f + 4 * b, *G = box_g + 4 * b;
        const long n1 = F[1] - F[0] + 1, n2 = F[3] - F[2] + 1;
        ok[b] = 0;
        out[3 * b] = out[3 * b + 1] = out[3 * b + 2] = NAN;
        const bool good = n1 >= 2 && n2 >= 2 && G[1] - G[0] + 1 == n1 && G[3] - G[2] + 1 == n2 &&
                          F[0] >= 1 && F[2] >= 1 && F[1] <= fr.H && F[3] <= fr.W &&
                          G[0] - pad >= 1 && G[2] - pad >= 1 && G[1] + pad <= fr.H &&
                          G[3] + pad <= fr.W;
        if (!good) continue;
        idx.push_back(b);
        bf.insert(bf.end(), F, F + 4);
        bg.insert(bg.end(), G, G + 4);
        tile0.push_back(tile0.back() + xcorr_box_tiles(n1, n2));
    }
    const long nb = (long)idx.size();
    if (nb == 0) return STAQ_OK;
    const long ntile = tile0.back();
    hipStream_t s = ctx->stream;
    HIPCHK(ctx, ctx->s[6].ensure((size_t)nb * 64 + (size_t)(nb + 1) * 4));
    HIPCHK(ctx, ctx->s[7].ensure((size_t)ntile * xcorr_box_part_doubles((int)pad) * 8));
    HIPCHK(ctx, ctx->s[8].ensure((size_t)nb * 24));
    HIPCHK(ctx, ctx->s[9].ensure((size_t)nb * 4));
    long *dbf = ctx->s[6].as<long>(), *dbg = dbf + 4 * nb;
    int *dt0 = reinterpret_cast<int *>(dbg + 4 * nb);
    HIPCHK(ctx, hipMemcpyAsync(dbf, bf.data(), nb * 32, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(dbg, bg.data(), nb * 32, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(dt0, tile0.data(), (nb + 1) * 4, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, xcorr_boxes_launch(fr, (int)pad, nb, ntile, dbf, dbg, dt0, ctx->s[7].as<double>(),
                                   ctx->s[8].as<double>(), ctx->s[9].as<int>(), s));
    std::vector<double> ho(3 * nb);
    std::vector<int> hk(nb);
    HIPCHK(ctx, hipMemcpyAsync(ho.data(), ctx->s[8].p, nb * 24, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipMemcpyAsync(hk.data(), ctx->s[9].p, nb * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    for (long k = 0; k < nb; ++k) {
        const long b = idx[k];
        ok[b] = hk[k];
        if (hk[k])
            for (int q = 0; q < 3; ++q) out[3 * b + q] = ho[3 * k + q];
    }
    return STAQ_OK;
}

int staq_xcorr_qfactors(staq_ctx *ctx, int slot, long nbox, const long *box_f, const long *box_g,
                        double *q, int *ok)
{
    if (!ctx || nbox < 0 || (nbox > 0 && (!box_f || !box_g || !q || !ok)))
        return fail(ctx, STAQ_EINVAL, "null argument");
    FrameDev fr;
    if (!frame_desc(ctx, slot, fr)) return fail(ctx, STAQ_ENOFRAME, "frame slot not set");
    if (nbox == 0) return STAQ_OK;
    hipSetDevice(ctx->device);
    // boxes that leave the image or whose g box is smaller than the template
    // are the reference's catch; the others are computed in chunks whose maps
    // fit the scratch budget
    constexpr size_t kMapBudget = (size_t)16 << 20;  // placements per chunk (256 MiB)
    constexpr size_t kLdsMax = 150 * 1024;
    std::vector<long> good;
    for (long b = 0; b < nbox; ++b) {
        const long *F = box_f + 4 * b, *G = box_g + 4 * b;
        const long n1 = F[1] - F[0] + 1, n2 = F[3] - F[2] + 1;
        const long d1 = G[1] - G[0] + 1, d2 = G[3] - G[2] + 1;
        ok[b] = 0;
        q[2 * b] = q[2 * b + 1] = NAN;
        const bool in = n1 >= 1 && n2 >= 1 && d1 >= n1 && d2 >= n2 && F[0] >= 1 && F[2] >= 1 &&
                        F[1] <= fr.H && F[3] <= fr.W && G[0] >= 1 && G[2] >= 1 && G[1] <= fr.H &&
                        G[3] <= fr.W;
        if (in && (size_t)(n1 + d1 - 1) * (size_t)(n2 + d2 - 1) > kMapBudget)
            return fail(ctx, STAQ_EINVAL, "a box's correlation map exceeds 16M placements");
        if (in) good.push_back(b);
    }
    hipStream_t s = ctx->stream;
    size_t pos = 0;
    while (pos < good.size()) {
        // one chunk: boxes in order while their maps fit
        std::vector<long> bf, bg, map0, lds_ix, hbm_ix;
        size_t np = 0, lds = 0;
        const size_t first = pos;
        while (pos < good.size()) {
            const long b = good[pos];
            const long *F = box_f + 4 * b, *G = box_g + 4 * b;
            const long n1 = F[1] - F[0] + 1, n2 = F[3] - F[2] + 1;
            const long d1 = G[1] - G[0] + 1, d2 = G[3] - G[2] + 1;
            const size_t nb = (size_t)(n1 + d1 - 1) * (size_t)(n2 + d2 - 1);
            if (np + nb > kMapBudget && pos > first) break;
            const long k = (long)(pos - first);
            bf.insert(bf.end(), F, F + 4);
            bg.insert(bg.end(), G, G + 4);
            map0.push_back((long)np);
            const size_t need = qfactor_lds_bytes(n1, n2, d1, d2);
            if (need <= kLdsMax) {
                lds_ix.push_back(k);
                lds = std::max(lds, need);
            } else {
                hbm_ix.push_back(k);
            }
            np += nb;
            ++pos;
        }
        const long nb = (long)(pos - first);
        const size_t ib = (size_t)nb * 10 * sizeof(long);
        HIPCHK(ctx, ctx->s[6].ensure(ib + (size_t)nb * (2 * sizeof(double) + sizeof(int))));
        HIPCHK(ctx, ctx->s[7].ensure(np * 2 * sizeof(double)));
        long *dbf = ctx->s[6].as<long>(), *dbg = dbf + 4 * nb, *dm0 = dbg + 4 * nb;
        long *dix = dm0 + nb;  // lds boxes, then hbm boxes
        double *dq = reinterpret_cast<double *>(dix + nb);
        int *dok = reinterpret_cast<int *>(dq + 2 * nb);
        std::vector<long> ix(lds_ix);
        ix.insert(ix.end(), hbm_ix.begin(), hbm_ix.end());
        HIPCHK(ctx, hipMemcpyAsync(dbf, bf.data(), nb * 32, hipMemcpyHostToDevice, s));
        HIPCHK(ctx, hipMemcpyAsync(dbg, bg.data(), nb * 32, hipMemcpyHostToDevice, s));
        HIPCHK(ctx, hipMemcpyAsync(dm0, map0.data(), nb * 8, hipMemcpyHostToDevice, s));
        HIPCHK(ctx, hipMemcpyAsync(dix, ix.data(), nb * 8, hipMemcpyHostToDevice, s));
        HIPCHK(ctx, qfactor_launch(fr, dbf, dbg, dm0, ctx->s[7].as<double>(), dq, dok, dix,
                                   (long)lds_ix.size(), lds, dix + lds_ix.size(), (long)hbm_ix.size(),
                                   s));
        std::vector<double> hq(2 * nb);
        std::vector<int> hk(nb);
        HIPCHK(ctx, hipMemcpyAsync(hq.data(), dq, nb * 16, hipMemcpyDeviceToHost, s));
        HIPCHK(ctx, hipMemcpyAsync(hk.data(), dok, nb * 4, hipMemcpyDeviceToHost, s));
        HIPCHK(ctx, hipStreamSynchronize(s));
        for (long k = 0; k < nb; ++k) {
            const long b = good[first + k];
            ok[b] = hk[k];
            if (hk[k]) {
                q[2 * b] = hq[2 * k];
                q[2 * b + 1] = hq[2 * k + 1];
            }
        }
    }
    return STAQ_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// batches
// ---------------------------------------------------------------------------
static int grown_window(int w)
{
    // upper bound of the window after the mask-driven growth:
    // 2*ceil(sqrt(n/(n-count))*w/2) with count < 0.4 n (funICGNQuadtree.m:70,
    // funICGN_Subpb1.m:78-79), or round(sqrt(2) w) (funICGN.m:58)
    const int a = 2 * (int)std::ceil(std::sqrt(1.0 / 0.6) * w / 2.0);
    const int b = (int)std::lround(std::sqrt(2.0) * w);
    return std::max(a, b) + 2;
}

static int tile_side(int n) { return (int)std::ceil(n * 1.12) + 8; }

// bins from this many subsets take one wavefront per subset up to 1152 pixels
static constexpr long kOneWaveMin = 2048;

static int default_bs(int npix)
{
    if (npix <= 400) return 64;
    if (npix <= 1300) return 128;
    return 256;
}

// Kernel forms of a bin: 0 / -1 = one wavefront per subset, -2 / -4 = two or
// four (icgn_wave.hip); 1 / 2 / 4 = the lean kernel with that many wavefronts,
// for windows without masked pixels (icgn_lean.hip); 64 .. 512 = one workgroup
// of that size per subset (icgn.hip).
static bool valid_form(int f)
{
    return f == 0 || f == -1 || f == -2 || f == -4 || f == 1 || f == 2 || f == 4 || f == 8 ||
           f == 16 || f == 64 || f == 128 || f == 256 || f == 512;
}
static bool lean_form(int f) { return f == 1 || f == 2 || f == 4 || f == 8 || f == 16; }
// pixel slots per lane of a wavefront form for windows of npix pixels (0: too large)
static int form_ppt(int f, int npix)
{
    if (lean_form(f)) return icgn_lean_ppt(npix, f);
    return f <= 0 ? icgn_wave_ppt(npix, f == 0 ? 1 : -f) : 0;
}
// wavefronts per subset of a wavefront form, 0 for a workgroup form
static int form_nw(int f)
{
    if (lean_form(f)) return f;
    return f <= 0 ? (f == 0 ? 1 : -f) : 0;
}


// Bins of similar window size; windows for the general kernel apart, grouped
// by the wavefront kernel their grown window needs (the growth rules of
// funICGNQuadtree.m:66-70 and funICGN_Subpb1.m:76-80); largest windows first,
// so that the hardware dispatcher fills each bin's tail with short subsets.
// (Re)builds the bins, their default forms, the side streams and events of
// the concurrent launches, the queue heads and the processing order.
static int build_bins(staq_batch *b)
{
    staq_ctx *ctx = b->ctx;
    const long N = b->N;
    const std::vector<int> &hx = b->hx, &hy = b->hy, &gnp = b->gnp, &gppt = b->gppt;
    const std::vector<char> &gen = b->general;
    std::vector<int> order(N);
    for (long i = 0; i < N; ++i) order[i] = (int)i;
    auto key = [&](int i) {
        return 64 * ((std::max(hx[i], hy[i]) + 7) / 8) + (gen[i] ? 32 + gppt[i] : 0);
    };
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return key(x) > key(y); });
    if (b->gexec) {  // a captured graph holds the old schedule
        hipGraphExecDestroy(b->gexec);
        hipGraphDestroy(b->graph);
        b->gexec = nullptr;
        b->graph = nullptr;
    }
    b->bins.clear();
    long i0 = 0;
    int fbx = 0, fby = 0;
    while (i0 < N) {
        const int k = key(order[i0]);
        long i1 = i0;
        Bin bin;
        while (i1 < N && key(order[i1]) == k) {
            bin.wxcap = std::max(bin.wxcap, hx[order[i1]]);
            bin.wycap = std::max(bin.wycap, hy[order[i1]]);
            bin.wave_npix = std::max(bin.wave_npix, gnp[order[i1]]);
            ++i1;
        }
        bin.item0 = i0;
        bin.n = i1 - i0;
        bin.masked = (k & 32) != 0;
        bin.npix_cap = (bin.wxcap + 1) * (bin.wycap + 1);
        bin.tile_cap = tile_side(bin.wxcap + 1) * tile_side(bin.wycap + 1);
        // the wave kernel of a bin that touches the mask covers the grown
        // windows of its smallest class too; a subset that outgrows its bin
        // (or grows although classified clear) defers to the fallback launch
        bin.wave_npix = std::max(bin.wave_npix, bin.npix_cap);
        bin.ppt = icgn_wave_ppt(bin.wave_npix, 1);
        // default form (measured on the hole, plain and large-deformation
        // workloads): bins without masked pixels take the lean kernel, one
        // wavefront per subset while five pixel slots hold the window, and
        // up to 1152 pixels (18 slots per lane: the 33 x 33 windows; no FP64
        // step taken twice) in bins of at least kOneWaveMin subsets, where
        // throughput decides; else two (shorter chains for small bins),
        // four or eight; bins for the general kernel take four (two, one)
        // wavefronts; windows beyond that take the workgroup kernel
        int form = default_bs(bin.wave_npix);
        const int g = bin.masked ? -1 : 1;
        if (!bin.masked && bin.ppt && bin.ppt <= 5)
            form = 1;
        else if (!bin.masked && bin.n >= kOneWaveMin && icgn_lean_ppt(bin.wave_npix, 1))
            form = 1;
        else if (!bin.masked && icgn_lean_ppt(bin.wave_npix, 2))
            form = 2;
        else if (icgn_wave_ppt(bin.wave_npix, 4))
            form = 4 * g;
        else if (icgn_wave_ppt(bin.wave_npix, 2))
            form = 2 * g;
        else if (bin.ppt)
            form = bin.masked ? 0 : 1;
        else if (!bin.masked && icgn_lean_ppt(bin.wave_npix, 8))
            form = 8;  // sixteen wavefronts spill registers (1024 threads cap them at 128)
        bin.bs[0] = bin.bs[1] = form;
        b->bins.push_back(bin);
        fbx = std::max(fbx, grown_window(bin.wxcap));
        fby = std::max(fby, grown_window(bin.wycap));
        i0 = i1;
    }
    b->fb_npix_cap = (fbx + 1) * (fby + 1);
    b->fb_tile_cap = tile_side(fbx + 1) * tile_side(fby + 1);
    const size_t nb = b->bins.size();
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    while (b->side.size() + 1 > std::max<size_t>(nb, 1)) {
        hipStreamSynchronize(b->side.back());
        hipStreamDestroy(b->side.back());
        b->side.pop_back();
    }
    while (b->side.size() + 1 < nb) {
        hipStream_t st = nullptr;
        HIPCHK(ctx, hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        b->side.push_back(st);
    }
    while (b->ev.size() > nb) {
        hipEventDestroy(b->ev.back());
        b->ev.pop_back();
    }
    while (b->ev.size() < nb) {
        hipEvent_t e = nullptr;
        HIPCHK(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming));
        b->ev.push_back(e);
    }
    HIPCHK(ctx, b->ints.ensure((nb + 5) * 4));
    HIPCHK(ctx, hipMemcpy(b->order.p, order.data(), N * 4, hipMemcpyHostToDevice));
    return STAQ_OK;
}

extern "C" {

int staq_batch_create(staq_ctx *ctx, staq_batch **out, long N, const double *coords,
                      const long *wx, const long *wy, const int *slot)
{
    if (!ctx || !out || N <= 0 || !coords || !wx) return fail(ctx, STAQ_EINVAL, "bad batch arguments");
    *out = nullptr;
    hipSetDevice(ctx->device);
    std::vector<int> hx(N), hy(N), hs(N, 0);
    int wmax = 0;
    for (long i = 0; i < N; ++i) {
        hx[i] = (int)wx[i];
        hy[i] = (int)(wy ? wy[i] : wx[i]);
        if (slot) hs[i] = slot[i];
        if (hx[i] < 2 || hy[i] < 2 || hx[i] > 2048 || hy[i] > 2048)
            return fail(ctx, STAQ_EINVAL, "window size out of range");
        if (hs[i] < 0 || (size_t)hs[i] >= ctx->frames.size() || !ctx->frames[hs[i]].set)
            return fail(ctx, STAQ_ENOFRAME, "subset refers to a frame slot that is not set");
        wmax = std::max(wmax, std::max(hx[i], hy[i]));
    }
    staq_batch *b = new staq_batch();
    b->ctx = ctx;
    b->N = N;
    b->wmax = wmax;
    b->has_slot = slot != nullptr;
    int rc = STAQ_OK;
    auto chk = [&](hipError_t e, const char *w) {
        if (e != hipSuccess && rc == STAQ_OK) rc = hipfail(ctx, e, w);
    };
    chk(b->coords.ensure(N * 16), "coords");
    chk(b->U0.ensure(N * 16), "U0");
    chk(b->UOld.ensure(N * 16), "UOld");
    chk(b->FOld.ensure(N * 32), "FOld");
    chk(b->vdual.ensure(N * 16), "vdual");
    chk(b->U.ensure(N * 16), "U");
    chk(b->F.ensure(N * 32), "F");
    chk(b->H.ensure(N * 21 * 8), "H");
    chk(b->status.ensure(N * 4), "status");
    chk(b->wx.ensure(N * 4), "wx");
    chk(b->wy.ensure(N * 4), "wy");
    chk(b->slot.ensure(N * 4), "slot");
    chk(b->order.ensure(N * 4), "order");
    // which windows touch the mask of their frame (their window grows at
    // set-up): classified on the device from the frames resident now
    std::vector<int> cnt(N, 0);
    if (rc == STAQ_OK) {
        hipStream_t s = ctx->stream;
        DevBuf dcls;
        chk(dcls.ensure(N * 4), "class");
        chk(hipMemcpyAsync(b->coords.p, coords, N * 16, hipMemcpyHostToDevice, s), "coords");
        chk(hipMemcpyAsync(b->wx.p, hx.data(), N * 4, hipMemcpyHostToDevice, s), "wx");
        chk(hipMemcpyAsync(b->wy.p, hy.data(), N * 4, hipMemcpyHostToDevice, s), "wy");
        chk(hipMemcpyAsync(b->slot.p, hs.data(), N * 4, hipMemcpyHostToDevice, s), "slot");
        if (rc == STAQ_OK)
            chk(mask_class_launch(ctx->d_frames.as<FrameDev>(), b->slot.as<int>(), b->coords.as<double>(),
                                  b->wx.as<int>(), b->wy.as<int>(), N, dcls.as<int>(), s),
                "mask classes");
        if (rc == STAQ_OK)
            chk(hipMemcpyAsync(cnt.data(), dcls.p, N * 4, hipMemcpyDeviceToHost, s), "class");
        chk(hipStreamSynchronize(s), "sync");
        dcls.release();
    }
    if (rc != STAQ_OK) {
        staq_batch_destroy(b);
        return rc;
    }
    // the window each subset's kernel must hold: windows that touch the mask
    // grow at set-up (funICGNQuadtree.m:66-70, funICGN_Subpb1.m:76-80)
    std::vector<int> gppt(N, 0), gnp(N, 0);
    for (long i = 0; i < N; ++i) {
        gnp[i] = (hx[i] + 1) * (hy[i] + 1);
        if (cnt[i] == 0) {
            const int pg = icgn_wave_ppt(gnp[i], 1);
            gppt[i] = pg ? pg : 31;
            continue;
        }
        const double n0 = (double)(hx[i] + 1) * (hy[i] + 1);
        int gx = hx[i], gy = hy[i];
        if (cnt[i] > 0 && cnt[i] < 0.4 * n0) {
            gx = 2 * (int)std::ceil(std::sqrt(n0 / (n0 - cnt[i])) * hx[i] / 2.0);
            const double n1 = (double)(gx + 1) * (hy[i] + 1);
            gy = std::max(gx, 2 * (int)std::ceil(std::sqrt(n1 / (n1 - cnt[i])) * hy[i] / 2.0));
            gx = std::max(gx, gy);
        }
        gnp[i] = (gx + 1) * (gy + 1);
        const int pg = icgn_wave_ppt(gnp[i], 1);
        gppt[i] = pg ? pg : 31;
    }
    b->hx = hx;
    b->hy = hy;
    b->cnt = cnt;
    b->gnp = gnp;
    b->gppt = gppt;
    b->general.assign(N, 0);
    for (long i = 0; i < N; ++i) b->general[i] = cnt[i] != 0;
    chk(b->handed.ensure(N * 4), "handed");
    chk(b->deflist.ensure(N * 4), "defer list");
    if (rc == STAQ_OK) rc = build_bins(b);
    if (rc == STAQ_OK) {
        hipStream_t s = ctx->stream;
        chk(hipMemsetAsync(b->handed.p, 0, N * 4, s), "zero");
        chk(hipMemsetAsync(b->U0.p, 0, N * 16, s), "zero");
        chk(hipMemsetAsync(b->UOld.p, 0, N * 16, s), "zero");
        chk(hipMemsetAsync(b->FOld.p, 0, N * 32, s), "zero");
        chk(hipMemsetAsync(b->vdual.p, 0, N * 16, s), "zero");
        chk(hipStreamSynchronize(s), "sync");
    }
    if (rc != STAQ_OK) {
        staq_batch_destroy(b);
        return rc;
    }
    *out = b;
    return STAQ_OK;
}

void staq_batch_destroy(staq_batch *b)
{
    if (!b) return;
    hipSetDevice(b->ctx->device);
    hipStreamSynchronize(b->ctx->stream);
    if (b->gexec) hipGraphExecDestroy(b->gexec);
    if (b->graph) hipGraphDestroy(b->graph);
    for (hipStream_t st : b->side) {
        hipStreamSynchronize(st);
        hipStreamDestroy(st);
    }
    for (hipEvent_t e : b->ev) hipEventDestroy(e);
    for (hipEvent_t e : b->tev) hipEventDestroy(e);
    DevBuf *all[] = {&b->coords, &b->U0, &b->UOld, &b->FOld, &b->vdual, &b->U, &b->F,
                     &b->H, &b->status, &b->wx, &b->wy, &b->slot, &b->order, &b->ints,
                     &b->handed, &b->deflist};
    for (DevBuf *d : all) d->release();
    delete b;
}

int staq_batch_set_input(staq_batch *b, int which, const double *host)
{
    if (!b || !host) return STAQ_EINVAL;
    staq_ctx *ctx = b->ctx;
    hipSetDevice(ctx->device);
    DevBuf *d = which == STAQ_IN_U0 ? &b->U0 : which == STAQ_IN_UOLD ? &b->UOld
              : which == STAQ_IN_FOLD ? &b->FOld : which == STAQ_IN_VDUAL ? &b->vdual : nullptr;
    if (!d) return fail(ctx, STAQ_EINVAL, "unknown input");
    const size_t n = (which == STAQ_IN_FOLD ? 4 : 2) * (size_t)b->N * 8;
    HIPCHK(ctx, hipMemcpyAsync(d->p, host, n, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return STAQ_OK;
}

int staq_batch_get_output(staq_batch *b, double *U, double *F, double *Hpar, int *status)
{
    if (!b) return STAQ_EINVAL;
    staq_ctx *ctx = b->ctx;
    hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    const size_t N = b->N;
    if (U) HIPCHK(ctx, hipMemcpyAsync(U, b->U.p, N * 16, hipMemcpyDeviceToHost, s));
    if (F) HIPCHK(ctx, hipMemcpyAsync(F, b->F.p, N * 32, hipMemcpyDeviceToHost, s));
    if (Hpar) HIPCHK(ctx, hipMemcpyAsync(Hpar, b->H.p, N * 21 * 8, hipMemcpyDeviceToHost, s));
    if (status) HIPCHK(ctx, hipMemcpyAsync(status, b->status.p, N * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    return STAQ_OK;
}

int staq_batch_exact_count(staq_batch *b, long *n)
{
    if (!b || !n) return STAQ_EINVAL;
    staq_ctx *ctx = b->ctx;
    hipSetDevice(ctx->device);
    int v = 0;
    const int nb = (int)b->bins.size();
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx, hipMemcpy(&v, b->ints.as<int>() + nb + 2, 4, hipMemcpyDeviceToHost));
    *n = v;
    return STAQ_OK;
}

int staq_batch_get_ptrs(staq_batch *b, staq_batch_ptrs *o)
{
    if (!b || !o) return STAQ_EINVAL;
    o->coords = b->coords.as<double>();
    o->U0 = b->U0.as<double>();
    o->UOld = b->UOld.as<double>();
    o->FOld = b->FOld.as<double>();
    o->vdual = b->vdual.as<double>();
    o->U = b->U.as<double>();
    o->F = b->F.as<double>();
    o->Hpar = b->H.as<double>();
    o->status = b->status.as<int>();
    o->wx = b->wx.as<int>();
    o->wy = b->wy.as<int>();
    o->slot = b->slot.as<int>();
    o->N = b->N;
    o->wmax = b->wmax;
    return STAQ_OK;
}

int staq_batch_bins(staq_batch *b, int *nbins, int *wcap, long *count, int *bs6, int *bs2,
                    int *ppt, int maxbins)
{
    if (!b || !nbins) return STAQ_EINVAL;
    *nbins = (int)b->bins.size();
    for (int k = 0; k < (int)b->bins.size() && k < maxbins; ++k) {
        if (wcap) wcap[k] = std::max(b->bins[k].wxcap, b->bins[k].wycap);
        if (count) count[k] = b->bins[k].n;
        if (bs6) bs6[k] = b->bins[k].bs[0];
        if (bs2) bs2[k] = b->bins[k].bs[1];
        if (ppt) {
            ppt[k] = form_ppt(b->bins[k].bs[0], b->bins[k].wave_npix);
        }
    }
    return STAQ_OK;
}

}  // extern "C"

extern "C" int staq_batch_set_form(staq_batch *b, int bin, int mode, int form)
{
    if (!b || bin < 0 || bin >= (int)b->bins.size() || (mode != 6 && mode != 2))
        return STAQ_EINVAL;
    if (!valid_form(form))
        return fail(b->ctx, STAQ_EINVAL,
                    "kernel form must be 0, -1, -2, -4, 1, 2, 4, 8, 16, 64, 128, 256 or 512");
    b->bins[bin].bs[mode == 6 ? 0 : 1] = form;
    if (b->gexec) {  // a captured graph holds the old schedule
        hipGraphExecDestroy(b->gexec);
        hipGraphDestroy(b->graph);
        b->gexec = nullptr;
        b->graph = nullptr;
    }
    return STAQ_OK;
}

extern "C" int staq_batch_bin_of(staq_batch *b, int *bin)
{
    if (!b || !bin) return STAQ_EINVAL;
    staq_ctx *ctx = b->ctx;
    hipSetDevice(ctx->device);
    std::vector<int> order(b->N);
    HIPCHK(ctx, hipMemcpy(order.data(), b->order.p, b->N * 4, hipMemcpyDeviceToHost));
    for (int k = 0; k < (int)b->bins.size(); ++k)
        for (long t = b->bins[k].item0; t < b->bins[k].item0 + b->bins[k].n; ++t) bin[order[t]] = k;
    return STAQ_OK;
}

// Stopping rule, damping, iteration cap and guard band of a local step.
// The guard bands as limits on the squared norms, with the operations of
// near_tol (common.h), so that the lean kernel's compares decide alike.
static void set_bands(IcgnArgs &a)
{
    const double tol2 = a.tol * a.tol;
    const bool r = a.guard2 > 0, b = a.guard2_abs > 0;
    a.rel_lo2 = r ? tol2 * (1.0 - a.guard2) : INFINITY;
    a.rel_hi2 = r ? tol2 * (1.0 + a.guard2) : -INFINITY;
    a.abs_lo2 = b ? tol2 * (1.0 - a.guard2_abs) : INFINITY;
    a.abs_hi2 = b ? tol2 * (1.0 + a.guard2_abs) : -INFINITY;
}

static void set_step_params(IcgnArgs &a, const staq_params *p)
{
    a.tol = p->tol;
    a.mu = p->mu;
    a.plain = p->plain;
    a.max_it = p->max_iter > 0 ? p->max_iter : 100;
    // the band on |b| / tol, (1 - g, 1 + g), as a band on the squared norms
    const double g = p->guard > 0 ? p->guard : 0.0;
    a.guard2 = g > 0 ? g * (2.0 + g) : 0.0;
    // |b| near tol is small in absolute terms, where the FP32 rounding is a
    // larger fraction of it: a band four times as wide
    const double ga = 4.0 * g;
    a.guard2_abs = ga > 0 ? ga * (2.0 + ga) : 0.0;
    set_bands(a);
}

static IcgnArgs base_args(staq_batch *b, const staq_params *p)
{
    IcgnArgs a;
    std::memset(&a, 0, sizeof(a));
    a.frames = b->ctx->d_frames.as<FrameDev>();
    a.slot = b->has_slot ? b->slot.as<int>() : nullptr;
    a.coords = b->coords.as<double>();
    a.wx = b->wx.as<int>();
    a.wy = b->wy.as<int>();
    a.order = b->order.as<int>();
    a.U0 = b->U0.as<double>();
    a.UOld = b->UOld.as<double>();
    a.FOld = b->FOld.as<double>();
    a.vdual = b->vdual.as<double>();
    a.U = b->U.as<double>();
    a.F = b->F.as<double>();
    a.H21 = b->H.as<double>();
    a.status = b->status.as<int>();
    a.handed = b->handed.as<int>();
    a.defer_list = b->deflist.as<int>();
    a.N = b->N;
    set_step_params(a, p);
    return a;
}

// The fallback launches over the defer list of a step: the FP64 re-solve of
// the subsets whose stopping test fell inside the guard band and of the
// windows that outgrew their kernel. Launched unconditionally (capturable in
// a graph); their workgroups leave at once when the list is empty. Two
// launches: windows of up to 32 px (LDS for their grown windows: g in FP64,
// DfDx and DfDy, the g tile; two subsets per CU), then the larger ones,
// gathering from HBM.
// q[0], q[1]: queue head and list length of the first; q[3], q[4]: of the
// second (q[2] counts the FP64 re-solves).  A subset is listed by its own
// window size, so with no window above kSmallFallbackW px (wmax) the second
// list stays empty and its launch is left out.
static hipError_t fallback_launches(IcgnArgs a, int mode, int *q, int fb_npix_cap, int wmax,
                                    hipStream_t s)
{
    a.item0 = 0;
    a.n_items = a.N;
    a.defer = 0;
    a.guard2 = 0;
    a.guard2_abs = 0;
    set_bands(a);
    a.only_deferred = 1;
    a.npix_cap = std::min(kSmallFallbackPix, fb_npix_cap);
    a.tile_cap = tile_side(48) * tile_side(48);  // the warped support of a 48 x 48 window
    a.grad_lds = 1;                              // DfDx, DfDy of the window in LDS
    a.work_counter = q;
    a.defer_flag = q + 1;
    a.list_back = 0;
    hipError_t e = icgn_exact_launch(a, mode, 256, (int)std::min<long>(a.N, 1280), s);
    if (e != hipSuccess || wmax <= kSmallFallbackW) return e;
    a.npix_cap = fb_npix_cap;
    a.tile_cap = 0;  // gathers from HBM
    a.grad_lds = 0;
    a.work_counter = q + 3;
    a.defer_flag = q + 4;
    a.list_back = 1;
    return icgn_exact_launch(a, mode, 256, (int)std::min<long>(a.N, 256), s);
}

static int occupancy_grid(int bs, size_t lds, long n)
{
    // persistent grid: enough workgroups to fill 256 CUs at the LDS-limited occupancy
    const size_t per_cu_lds = 160 * 1024;
    int per_cu = (int)std::max<size_t>(1, per_cu_lds / std::max<size_t>(lds + 4096, 1));
    per_cu = std::min(per_cu, std::max(1, 2048 / bs));
    const long g = (long)per_cu * 256;
    return (int)std::min<long>(n, g);
}

// Enqueue one local step (mode 6 or 2) on stream s: reset of the queue heads
// and defer flag, the window bins (concurrently when side streams exist), and
// the fallback for windows that outgrew their bin.  bin_only >= 0 launches a
// single bin on s (autotuning), without fallback.
static constexpr int kNoOverride = 1 << 30;

static int enqueue_step(staq_batch *b, const staq_params *p, int mode, hipStream_t s,
                        int bin_only = -1, int bs_override = kNoOverride, bool timed = false)
{
    staq_ctx *ctx = b->ctx;
    const int nb = (int)b->bins.size();
    // [0, nb) bin queue heads; nb, nb+1: queue head and list length of the small
    // fallback launch; nb+2: count of subsets left to the FP64 re-solve; nb+3,
    // nb+4: queue head and list length of the large fallback launch
    int *ints = b->ints.as<int>();
    hipLaunchKernelGGL(reset_ints_kernel, dim3(1), dim3(256), 0, s, ints, nb + 5);
    HIPCHK(ctx, hipGetLastError());
    const int mi = mode == 6 ? 0 : 1;
    const bool fork = bin_only < 0 && nb > 1 && (int)b->side.size() == nb - 1 &&
                      (int)b->ev.size() == nb;
    if (fork) HIPCHK(ctx, hipEventRecord(b->ev[0], s));
    for (int k = 0; k < nb; ++k) {
        if (bin_only >= 0 && k != bin_only) continue;
        hipStream_t sk = s;
        if (fork && k > 0) {
            sk = b->side[k - 1];
            HIPCHK(ctx, hipStreamWaitEvent(sk, b->ev[0], 0));
        }
        const Bin &bin = b->bins[k];
        IcgnArgs a = base_args(b, p);
        a.item0 = bin.item0;
        a.n_items = bin.n;
        a.npix_cap = bin.npix_cap;
        a.tile_cap = bin.tile_cap;
        a.defer = 1;
        a.defer_out = ints + nb + 1;
        a.defer_out2 = ints + nb + 4;
        a.exact_count = ints + nb + 2;
        int bs = bs_override != kNoOverride ? bs_override
                                             : (p->block_size ? p->block_size : bin.bs[mi]);
        if (!valid_form(bs))
            return fail(ctx, STAQ_EINVAL,
                        "block_size must be 0, -1, -2, -4, 1, 2, 4, 8, 16, 64, 128, 256 or 512");
        const int nw = form_nw(bs);
        const bool lean = lean_form(bs);
        const int wppt = nw ? form_ppt(bs, bin.wave_npix) : 0;
        if (nw && !wppt) bs = default_bs(bin.wave_npix);
        int grid = (int)bin.n;
        if (timed) HIPCHK(ctx, hipEventRecord(b->tev[2 * k], sk));
        if (nw && wppt) {
            // nw wavefronts per subset; the g tile holds the warped window
            // of the largest window the kernel takes, plus its margin
            const int side = tile_side((int)std::sqrt(64.0 * wppt * nw));
            a.tile_cap = side * side;
            if (p->persistent) {
                a.work_counter = ints + k;
                // the g tile, f of every pixel slot, the exchange buffers
                grid = occupancy_grid(64 * nw, sizeof(float) * (a.tile_cap + 64 * wppt * nw) + 8192,
                                      bin.n);
            }
            if (lean)
                HIPCHK(ctx, icgn_lean_launch(a, mode, wppt, nw, grid, sk));
            else
                HIPCHK(ctx, icgn_wave_launch(a, mode, wppt, nw, grid, sk));
        } else {
            // the workgroup kernel holds the windows its subsets grow to at
            // set-up, so none of them defers to the fallback launch
            const int gside = tile_side((int)std::ceil(std::sqrt((double)bin.wave_npix)));
            a.npix_cap = std::max(bin.npix_cap, bin.wave_npix);
            a.tile_cap = std::max(bin.tile_cap, gside * gside);
            if (p->persistent) {
                a.work_counter = ints + k;
                grid = occupancy_grid(bs, icgn_lds_bytes(a.npix_cap, a.tile_cap), bin.n);
            }
            HIPCHK(ctx, icgn_launch(a, mode, bs, grid, sk));
        }
        if (timed) HIPCHK(ctx, hipEventRecord(b->tev[2 * k + 1], sk));
        if (fork && k > 0) {
            HIPCHK(ctx, hipEventRecord(b->ev[k], sk));
            HIPCHK(ctx, hipStreamWaitEvent(s, b->ev[k], 0));
        }
    }
    if (bin_only < 0) {
        // fallback over the whole batch
        if (timed) HIPCHK(ctx, hipEventRecord(b->tev[2 * nb], s));
        HIPCHK(ctx, fallback_launches(base_args(b, p), mode, ints + nb, b->fb_npix_cap, b->wmax, s));
        if (timed) HIPCHK(ctx, hipEventRecord(b->tev[2 * nb + 1], s));
    }
    return STAQ_OK;
}

extern "C" {

int staq_batch_icgn6(staq_batch *b, const staq_params *p)
{
    if (!b || !p) return STAQ_EINVAL;
    hipSetDevice(b->ctx->device);
    return enqueue_step(b, p, 6, b->ctx->stream);
}

int staq_batch_icgn2(staq_batch *b, const staq_params *p)
{
    if (!b || !p) return STAQ_EINVAL;
    hipSetDevice(b->ctx->device);
    return enqueue_step(b, p, 2, b->ctx->stream);
}

int staq_batch_step_timed(staq_batch *b, const staq_params *p, int mode, float *kernel_ms)
{
    if (!b || !p || !kernel_ms || (mode != 6 && mode != 2)) return STAQ_EINVAL;
    staq_ctx *ctx = b->ctx;
    hipSetDevice(ctx->device);
    const int nb = (int)b->bins.size();
    while ((int)b->tev.size() < 2 * nb + 2) {
        hipEvent_t e;
        HIPCHK(ctx, hipEventCreate(&e));
        b->tev.push_back(e);
    }
    int rc = enqueue_step(b, p, mode, ctx->stream, -1, kNoOverride, true);
    if (rc != STAQ_OK) return rc;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    for (int k = 0; k <= nb; ++k)
        HIPCHK(ctx, hipEventElapsedTime(&kernel_ms[k], b->tev[2 * k], b->tev[2 * k + 1]));
    return STAQ_OK;
}

int staq_batch_icgn6_per_subset(staq_batch *b, const staq_params *p, double *kernel_ms_sum)
{
    if (!b || !p) return STAQ_EINVAL;
    staq_ctx *ctx = b->ctx;
    hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    // Every subset gets the largest caps so that no subset defers: this is the
    // reference's one-subset-per-call structure with one workgroup per call.
    const int npix_cap = b->fb_npix_cap, tile_cap = b->fb_tile_cap;
    const int bs = p->block_size >= 64 ? p->block_size : 256;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (kernel_ms_sum) {
        *kernel_ms_sum = 0.0;
        HIPCHK(ctx, hipEventCreate(&e0));
        HIPCHK(ctx, hipEventCreate(&e1));
    }
    for (long k = 0; k < b->N; ++k) {
        IcgnArgs a = base_args(b, p);
        a.item0 = k;
        a.n_items = 1;
        a.npix_cap = npix_cap;
        a.tile_cap = tile_cap;
        a.guard2 = a.guard2_abs = 0;  // no fallback launch follows: every subset is finished here
        set_bands(a);
        if (kernel_ms_sum) HIPCHK(ctx, hipEventRecord(e0, s));
        HIPCHK(ctx, icgn_launch(a, 6, bs, 1, s));
        if (kernel_ms_sum) {
            HIPCHK(ctx, hipEventRecord(e1, s));
            HIPCHK(ctx, hipEventSynchronize(e1));
            float ms = 0.f;
            HIPCHK(ctx, hipEventElapsedTime(&ms, e0, e1));
            *kernel_ms_sum += ms;
        }
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    return STAQ_OK;
}

int staq_batch_reorder(staq_batch *b)
{
    // Subsets the lean kernel handed to the fallback launch in the last step
    // (a zero of the warped g, degenerate statistics) move to the general
    // kernel's bins, so that later steps solve them in their bin instead of
    // after every bin.  Then longest-processing-time-first order inside each
    // bin, predicted from the last step: iterations used (101 for a subset
    // that ran out of iterations) times window pixels.  Results do not
    // depend on the order.
    if (!b) return STAQ_EINVAL;
    staq_ctx *ctx = b->ctx;
    hipSetDevice(ctx->device);
    const long N = b->N;
    {
        std::vector<int> hd(N);
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        HIPCHK(ctx, hipMemcpy(hd.data(), b->handed.p, N * 4, hipMemcpyDeviceToHost));
        bool moved = false;
        for (long i = 0; i < N; ++i)
            if (hd[i] && !b->general[i]) {
                b->general[i] = 1;
                moved = true;
            }
        if (moved) {
            const int rc = build_bins(b);
            if (rc != STAQ_OK) return rc;
            HIPCHK(ctx, hipMemset(b->handed.p, 0, N * 4));
        }
    }
    std::vector<int> st(N), order(N), wx(N), wy(N);
    HIPCHK(ctx, hipMemcpy(st.data(), b->status.p, N * 4, hipMemcpyDeviceToHost));
    HIPCHK(ctx, hipMemcpy(order.data(), b->order.p, N * 4, hipMemcpyDeviceToHost));
    HIPCHK(ctx, hipMemcpy(wx.data(), b->wx.p, N * 4, hipMemcpyDeviceToHost));
    HIPCHK(ctx, hipMemcpy(wy.data(), b->wy.p, N * 4, hipMemcpyDeviceToHost));
    auto cost = [&](int i) {
        const int s = st[i];
        const double it = (s >= 1 && s <= 101) ? s : (s == ST_DEFER ? 50 : 1);
        return it * (wx[i] + 1.0) * (wy[i] + 1.0);
    };
    for (const Bin &bin : b->bins)
        std::stable_sort(order.begin() + bin.item0, order.begin() + bin.item0 + bin.n,
                         [&](int x, int y) { return cost(x) > cost(y); });
    // a captured graph reads the order array when it runs, so it stays valid
    HIPCHK(ctx, hipMemcpy(b->order.p, order.data(), N * 4, hipMemcpyHostToDevice));
    return STAQ_OK;
}

int staq_batch_chain_outputs(staq_batch *b)
{
    if (!b) return STAQ_EINVAL;
    hipSetDevice(b->ctx->device);
    const unsigned g = (unsigned)((4 * b->N + 255) / 256);
    hipLaunchKernelGGL(chain_kernel, dim3(g), dim3(256), 0, b->ctx->stream, b->U.as<double>(),
                       b->F.as<double>(), b->UOld.as<double>(), b->FOld.as<double>(),
                       b->vdual.as<double>(), b->N);
    HIPCHK(b->ctx, hipGetLastError());
    return STAQ_OK;
}

int staq_batch_autotune(staq_batch *b, const staq_params *p, int mode)
{
    // Coordinate descent over the window bins: for each bin try the kernel
    // forms (one, two or four wavefronts per subset, each workgroup size).
    // The bins of a step run concurrently, so a step lasts as long as its
    // longest launch: each form is judged by the longest launch of an eager
    // step, every launch timed by events on its own stream (the host's launch
    // gaps do not enter), median of three steps; a form replaces the current
    // one only if it is faster by more than the timing noise.
    if (!b || !p || (mode != 6 && mode != 2)) return STAQ_EINVAL;
    staq_ctx *ctx = b->ctx;
    hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    const int nb = (int)b->bins.size();
    while ((int)b->tev.size() < 2 * nb + 2) {
        hipEvent_t e;
        HIPCHK(ctx, hipEventCreate(&e));
        b->tev.push_back(e);
    }
    const int cand[12] = {0, -2, -4, 1, 2, 4, 8, 16, 64, 128, 256, 512};
    const int mi = mode == 6 ? 0 : 1;
    staq_params q = *p;
    q.block_size = 0;
    int rc = STAQ_OK;
    auto time_step = [&](float *ms) {
        float t[3] = {0, 0, 0};
        for (int r = 0; r < 3 && rc == STAQ_OK; ++r) {
            rc = enqueue_step(b, &q, mode, s, -1, kNoOverride, true);
            if (rc != STAQ_OK) break;
            hipError_t e = hipStreamSynchronize(s);
            for (int k = 0; k <= nb && e == hipSuccess; ++k) {
                float d = 0.f;
                e = hipEventElapsedTime(&d, b->tev[2 * k], b->tev[2 * k + 1]);
                t[r] = std::max(t[r], d);
            }
            if (e != hipSuccess) rc = hipfail(ctx, e, "autotune timing");
        }
        std::sort(t, t + 3);
        *ms = t[1];
    };
    for (int k = 0; k < (int)b->bins.size() && rc == STAQ_OK; ++k) {
        Bin &bin = b->bins[k];
        const int cur = bin.bs[mi];
        float best = 0.f;
        time_step(&best);
        int best_bs = cur;
        for (int c : cand) {
            if (c == cur) continue;
            if (form_nw(c) && !form_ppt(c, bin.wave_npix)) continue;
            if (lean_form(c) && bin.masked) continue;  // its subsets would all defer
            if (c == 512 && bin.npix_cap < 1500) continue;
            if (c == 64 && bin.npix_cap > 4500) continue;
            bin.bs[mi] = c;
            float ms = 0.f;
            time_step(&ms);
            if (rc != STAQ_OK) break;
            if (ms < 0.97f * best) {
                best = ms;
                best_bs = c;
            }
        }
        bin.bs[mi] = best_bs;
    }
    if (b->gexec) {  // a captured graph holds the old schedule
        hipGraphExecDestroy(b->gexec);
        hipGraphDestroy(b->graph);
        b->gexec = nullptr;
        b->graph = nullptr;
    }
    return rc;
}

int staq_batch_graph_capture(staq_batch *b, const staq_params *p, int mode)
{
    if (!b || !p || (mode != 6 && mode != 2)) return STAQ_EINVAL;
    staq_ctx *ctx = b->ctx;
    hipSetDevice(ctx->device);
    if (b->gexec) {
        hipGraphExecDestroy(b->gexec);
        hipGraphDestroy(b->graph);
        b->gexec = nullptr;
        b->graph = nullptr;
    }
    hipStream_t s = ctx->stream;
    HIPCHK(ctx, hipStreamSynchronize(s));
    HIPCHK(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_step(b, p, mode, s);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (rc != STAQ_OK) {
        if (g) hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hipfail(ctx, e, "hipStreamEndCapture");
    hipGraphExec_t ge = nullptr;
    const hipError_t e2 = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    if (e2 != hipSuccess) {
        hipGraphDestroy(g);
        return hipfail(ctx, e2, "hipGraphInstantiate");
    }
    b->graph = g;
    b->gexec = ge;
    b->graph_mode = mode;
    return STAQ_OK;
}

int staq_batch_graph_launch(staq_batch *b)
{
    if (!b || !b->gexec) return fail(b ? b->ctx : nullptr, STAQ_EINVAL, "no captured graph");
    hipSetDevice(b->ctx->device);
    HIPCHK(b->ctx, hipGraphLaunch(b->gexec, b->ctx->stream));
    return STAQ_OK;
}

// ---------------------------------------------------------------------------
// host-pointer drop-ins built on a temporary batch
// ---------------------------------------------------------------------------
int staq_local_icgn(staq_ctx *ctx, const staq_params *p, long N, const double *U0,
                    const double *coords, const long *winsize, const int *slot, double *U,
                    double *F, double *Hpar, int *status)
{
    if (!ctx || !p || N <= 0 || !U0 || !coords || !winsize)
        return fail(ctx, STAQ_EINVAL, "null argument");
    staq_batch *b = nullptr;
    int rc = staq_batch_create(ctx, &b, N, coords, winsize, nullptr, slot);
    if (rc) return rc;
    rc = staq_batch_set_input(b, STAQ_IN_U0, U0);
    if (!rc) rc = staq_batch_icgn6(b, p);
    if (!rc) rc = staq_batch_get_output(b, U, F, Hpar, status);
    staq_batch_destroy(b);
    return rc;
}

int staq_subpb1(staq_ctx *ctx, const staq_params *p, long N, const double *UOld,
                const double *FOld, const double *udual, const double *vdual,
                const double *coords, const long *wx, const long *wy, const int *slot,
                double *U, int *status)
{
    (void)udual;
    if (!ctx || !p || N <= 0 || !UOld || !FOld || !vdual || !coords || !wx)
        return fail(ctx, STAQ_EINVAL, "null argument");
    staq_batch *b = nullptr;
    int rc = staq_batch_create(ctx, &b, N, coords, wx, wy, slot);
    if (rc) return rc;
    rc = staq_batch_set_input(b, STAQ_IN_UOLD, UOld);
    if (!rc) rc = staq_batch_set_input(b, STAQ_IN_FOLD, FOld);
    if (!rc) rc = staq_batch_set_input(b, STAQ_IN_VDUAL, vdual);
    if (!rc) rc = staq_batch_icgn2(b, p);
    if (!rc) rc = staq_batch_get_output(b, U, nullptr, nullptr, status);
    staq_batch_destroy(b);
    return rc;
}

// ---------------------------------------------------------------------------
// the 6-DOF local step without device allocation
// ---------------------------------------------------------------------------
int staq_local_icgn_noalloc(staq_ctx *ctx, const staq_params *p, int slot, long N, int wmax,
                            const double *U0, const double *coords, const int *winsize,
                            double *U, double *F, double *Hpar, int *status, void *stream)
{
    if (!ctx || !p || !U0 || !coords || !winsize || !U || !F || !Hpar || !status)
        return fail(ctx, STAQ_EINVAL, "null argument");
    if (N <= 0 || N > kNoAllocMaxN)
        return fail(ctx, STAQ_EINVAL, "staq_local_icgn_noalloc takes 1 to 1048576 subsets per call");
    if (wmax < 2 || wmax > 64)
        return fail(ctx, STAQ_EINVAL, "staq_local_icgn_noalloc takes windows of 2 to 64 px");
    if (slot < 0 || (size_t)slot >= ctx->frames.size() || !ctx->frames[slot].set)
        return fail(ctx, STAQ_ENOFRAME, "frame slot not set");
    if (!ctx->noalloc) return fail(ctx, STAQ_EHIP, "no device memory for the no-alloc entry");
    hipSetDevice(ctx->device);
    hipStream_t s = pick_stream(ctx, stream);
    int *q = ctx->noalloc;  // counters, then the defer list
    IcgnArgs a;
    std::memset(&a, 0, sizeof(a));
    a.frames = ctx->d_frames.as<FrameDev>() + slot;  // every subset on this slot
    a.coords = coords;
    a.wx = a.wy = winsize;
    a.U0 = U0;
    a.U = U;
    a.F = F;
    a.H21 = Hpar;
    a.status = status;
    a.defer_list = q + kNoAllocHead;
    a.N = N;
    a.item0 = 0;
    a.n_items = N;
    a.defer = 1;
    a.defer_out = q + 1;
    a.defer_out2 = q + 4;
    a.exact_count = q + 2;
    set_step_params(a, p);
    // the largest window after the mask-driven growth
    const int gw = grown_window(wmax);
    const int gnpix = (gw + 1) * (gw + 1);
    hipLaunchKernelGGL(reset_ints_kernel, dim3(1), dim3(64), 0, s, q, 5);
    HIPCHK(ctx, hipGetLastError());
    const int ppt = icgn_wave_ppt(gnpix, 4);
    if (ppt) {
        // four wavefronts per subset; windows that outgrow them go to the fallback
        const int side = tile_side((int)std::sqrt(64.0 * ppt * 4));
        a.tile_cap = side * side;
        HIPCHK(ctx, icgn_wave_launch(a, 6, ppt, 4, (int)N, s));
    } else {
        // the workgroup kernel holds the grown windows
        a.npix_cap = gnpix;
        a.tile_cap = tile_side(gw + 1) * tile_side(gw + 1);
        HIPCHK(ctx, icgn_launch(a, 6, 256, (int)N, s));
    }
    HIPCHK(ctx, fallback_launches(a, 6, q, gnpix, wmax, s));
    return STAQ_OK;
}

// ---------------------------------------------------------------------------
// ZNSSD forward / backward
// ---------------------------------------------------------------------------
static int znssd_dev(staq_ctx *ctx, long N, int wmax, const double *P, const double *coords,
                     const int *winsize, const int *slot, const double *gcost, double *cost,
                     double *gradP, int *status, hipStream_t s)
{
    if (!ctx || N <= 0 || wmax < 2 || !P || !coords || !winsize)
        return fail(ctx, STAQ_EINVAL, "bad znssd arguments");
    if (ctx->frames.empty()) return fail(ctx, STAQ_ENOFRAME, "no frame set");
    const long np = (long)(wmax + 1) * (wmax + 1);
    if (np * 12 > 150 * 1024) return fail(ctx, STAQ_EINVAL, "window too large for the adjoint kernel");
    ZnssdArgs a;
    a.frames = ctx->d_frames.as<FrameDev>();
    a.slot = slot;
    a.coords = coords;
    a.winsize = winsize;
    a.P = P;
    a.gcost = gcost;
    a.cost = cost;
    a.gradP = gradP;
    a.status = status;
    a.N = N;
    a.npix_cap = (int)np;
    HIPCHK(ctx, znssd_launch(a, s));
    return STAQ_OK;
}

int staq_znssd_fwd_dev(staq_ctx *ctx, long N, int wmax, const double *P, const double *coords,
                       const int *winsize, const int *slot, double *cost, int *status,
                       void *stream)
{
    if (!ctx) return STAQ_EINVAL;
    hipSetDevice(ctx->device);
    return znssd_dev(ctx, N, wmax, P, coords, winsize, slot, nullptr, cost, nullptr, status,
                     pick_stream(ctx, stream));
}

int staq_znssd_bwd_dev(staq_ctx *ctx, long N, int wmax, const double *P, const double *coords,
                       const int *winsize, const int *slot, const double *gcost, double *gradP,
                       void *stream)
{
    if (!ctx) return STAQ_EINVAL;
    hipSetDevice(ctx->device);
    return znssd_dev(ctx, N, wmax, P, coords, winsize, slot, gcost, nullptr, gradP, nullptr,
                     pick_stream(ctx, stream));
}

static int znssd_host(staq_ctx *ctx, long N, const double *P, const double *coords,
                      const long *winsize, const int *slot, const double *gcost, double *cost,
                      int *status, double *gradP, double *gradU)
{
    if (!ctx || N <= 0 || !P || !coords || !winsize) return fail(ctx, STAQ_EINVAL, "null argument");
    hipSetDevice(ctx->device);
    std::vector<int> w(N), sl(slot ? N : 0);
    int wmax = 2;
    for (long i = 0; i < N; ++i) {
        w[i] = (int)winsize[i];
        wmax = std::max(wmax, w[i]);
        if (slot) {
            sl[i] = slot[i];
            if (sl[i] < 0 || (size_t)sl[i] >= ctx->frames.size() || !ctx->frames[sl[i]].set)
                return fail(ctx, STAQ_ENOFRAME, "subset refers to a frame slot that is not set");
        }
    }
    hipStream_t s = ctx->stream;
    HIPCHK(ctx, ctx->s[0].ensure(N * 48));
    HIPCHK(ctx, ctx->s[1].ensure(N * 16));
    HIPCHK(ctx, ctx->s[2].ensure(N * 4));
    HIPCHK(ctx, ctx->s[3].ensure(N * 4));
    HIPCHK(ctx, ctx->s[4].ensure(N * 8));
    HIPCHK(ctx, ctx->s[5].ensure(N * 8));
    HIPCHK(ctx, ctx->s[6].ensure(N * 48));
    HIPCHK(ctx, ctx->s[7].ensure(N * 4));
    HIPCHK(ctx, hipMemcpyAsync(ctx->s[0].p, P, N * 48, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ctx->s[1].p, coords, N * 16, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ctx->s[2].p, w.data(), N * 4, hipMemcpyHostToDevice, s));
    if (slot) HIPCHK(ctx, hipMemcpyAsync(ctx->s[3].p, sl.data(), N * 4, hipMemcpyHostToDevice, s));
    if (gcost) HIPCHK(ctx, hipMemcpyAsync(ctx->s[4].p, gcost, N * 8, hipMemcpyHostToDevice, s));
    const int *dsl = slot ? ctx->s[3].as<int>() : nullptr;
    int rc;
    if (cost || status) {
        rc = znssd_dev(ctx, N, wmax, ctx->s[0].as<double>(), ctx->s[1].as<double>(),
                       ctx->s[2].as<int>(), dsl, nullptr, ctx->s[5].as<double>(), nullptr,
                       ctx->s[7].as<int>(), s);
        if (rc) return rc;
        if (cost) HIPCHK(ctx, hipMemcpyAsync(cost, ctx->s[5].p, N * 8, hipMemcpyDeviceToHost, s));
        if (status) HIPCHK(ctx, hipMemcpyAsync(status, ctx->s[7].p, N * 4, hipMemcpyDeviceToHost, s));
    }
    if (gradP || gradU) {
        rc = znssd_dev(ctx, N, wmax, ctx->s[0].as<double>(), ctx->s[1].as<double>(),
                       ctx->s[2].as<int>(), dsl, gcost ? ctx->s[4].as<double>() : nullptr, nullptr,
                       ctx->s[6].as<double>(), nullptr, s);
        if (rc) return rc;
        std::vector<double> gp(6 * N);
        HIPCHK(ctx, hipMemcpyAsync(gp.data(), ctx->s[6].p, N * 48, hipMemcpyDeviceToHost, s));
        HIPCHK(ctx, hipStreamSynchronize(s));
        if (gradP) std::memcpy(gradP, gp.data(), N * 48);
        if (gradU)
            for (long i = 0; i < N; ++i) {
                gradU[2 * i] = gp[6 * i + 4];
                gradU[2 * i + 1] = gp[6 * i + 5];
            }
    }
    HIPCHK(ctx, hipStreamSynchronize(s));
    return STAQ_OK;
}

int staq_znssd_fwd(staq_ctx *ctx, long N, const double *P, const double *coords,
                   const long *winsize, const int *slot, double *cost, int *status)
{
    return znssd_host(ctx, N, P, coords, winsize, slot, nullptr, cost, status, nullptr, nullptr);
}

int staq_znssd_bwd(staq_ctx *ctx, long N, const double *P, const double *coords,
                   const long *winsize, const int *slot, const double *gcost, double *gradP,
                   double *gradU)
{
    return znssd_host(ctx, N, P, coords, winsize, slot, gcost, nullptr, nullptr, gradP, gradU);
}

}  // extern "C"
