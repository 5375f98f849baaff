// rtm_capi.hip -- implementation of the C ABI declared in include/rtm_hip.h.
//
// One library stream per process; every operator validates its arguments on
// the host (shapes against what the kernels assume), enqueues its kernels
// and returns.  The time-loop entry points chain the operators of the
// reference's per-step loops with no host synchronisation in between.
#include "rtm_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rtm_hip.h"

using namespace rtmhip;

// Staging array of rtmhip_batch_steps_noalloc: the shot table of one call
// (at most kNoAllocShots entries, 64 x sizeof(BatchShot) = about 12 KB).
// Every GPU the library runs on holds its own copy until the process ends.
// Host code never passes `scratch` itself to a kernel or a copy: its device
// address on each GPU is looked up once, at rtmhip_init with that GPU
// current, and kept in g_scratch.
constexpr int kNoAllocShots = 64;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) BatchShot scratch[kNoAllocShots];

namespace {

BatchShot *g_scratch[kMaxDevices] = {};  // device address of `scratch` per GPU
std::mutex g_noalloc_mu;                  // guards the no-allocation host tables

struct Timing {
    std::vector<hipEvent_t> start, stop;
    size_t used = 0;
};

struct State {
    bool init = false;
    int device = -1;
    hipStream_t stream = nullptr;
    int *flag = nullptr;  // NaN ("numerical dispersion") flag written by the stencil
    bool timing = false;
    Timing tm[4];
    StencilTiling tiling[4] = {};  // per kernel class; tz == 0: built-in default
    int tiling_gen = 0;            // bumped by every rtmhip_set_tiling call
    // per wave format: the HBC upper tapers of the single-shot loops applied
    // in place by a separate pass (true) or fused into the stencil (false);
    // the modeling loop's every-border tapers fused (true) or separate
    bool taper_sep[3] = {false, true, true};
    int taper_all_mode[3] = {2, 2, 2};  // modeling taper3D: 0 separate, 1 on load, 2 on output
    int taper_mode[3] = {-1, -1, -1};   // the last rtmhip_set_taper_fusion mode
    bool taper_outu[3] = {true, false, false};  // forward upper tapers on the output
};
constexpr bool kTaperSepDefault[3] = {false, true, true};
// forward upper tapers on the output by default for F32; for 16-bit fields
// the separate pass before the two-z step is as fast (salt grid, 300 forward
// steps: FP32 0.136 s against 0.154 fused on load; FP16 0.110 against 0.107)
constexpr bool kTaperOutDefault[3] = {true, false, false};

State g_state;
thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int hip_fail(hipError_t e, const char *what)
{
    return fail(RTMHIP_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

#define HIPCHK(expr, what)                              \
    do {                                                \
        hipError_t e__ = (expr);                        \
        if (e__ != hipSuccess) return hip_fail(e__, what); \
    } while (0)

#define NEED_INIT()                                                               \
    do {                                                                          \
        if (!g_state.init) return fail(RTMHIP_ERR_NOT_INIT, "rtmhip_init() not called"); \
    } while (0)

int check_grid(int nxe, int nye, int nze)
{
    if (nxe <= 0 || nye <= 0 || nze <= 0)
        return fail(RTMHIP_ERR_INVALID, "grid dimensions must be positive (%d,%d,%d)", nxe, nye,
                    nze);
    if ((long long)nye * nze >= (1LL << 31))
        return fail(RTMHIP_ERR_INVALID, "nye*nze must stay below 2^31");
    return RTMHIP_OK;
}

int check_order(int order, int *radius)
{
    if (order != 2 && order != 4 && order != 6 && order != 8)
        return fail(RTMHIP_ERR_UNSUPPORTED, "stencil order %d not supported (2, 4, 6, 8)", order);
    *radius = order / 2;
    return RTMHIP_OK;
}

Coefs make_coefs(const float *c, int order)
{
    Coefs cf;
    memset(&cf, 0, sizeof cf);
    for (int d = 0; d < 3; ++d)
        for (int k = 0; k <= order; ++k) cf.c[d][k] = c[d * (order + 1) + k];
    return cf;
}

int check_rcv(const rtmhip_rcv_geom *g, int nze, int blen)
{
    if (!g) return fail(RTMHIP_ERR_INVALID, "receiver geometry is NULL");
    if (g->count_x < 0 || g->count_y < 0)
        return fail(RTMHIP_ERR_INVALID, "negative receiver count");
    if ((g->count_x > 1 && g->dist_x < 1) || (g->count_y > 1 && g->dist_y < 1))
        return fail(RTMHIP_ERR_INVALID,
                    "receiver distances must be >= 1 when more than one receiver per line");
    if (g->depth_z < 0 || g->depth_z + blen >= nze)
        return fail(RTMHIP_ERR_INVALID, "receiver depth %d outside the grid", g->depth_z);
    if (g->start_x < 0 || g->start_y < 0)
        return fail(RTMHIP_ERR_INVALID, "negative receiver start");
    return RTMHIP_OK;
}

// Runs launch() on the library stream, bracketed by HIP events of timing
// class cls when timing is on.
template <typename L>
hipError_t timed(int cls, L &&launch)
{
    Timing *tm = g_state.timing ? &g_state.tm[cls] : nullptr;
    if (tm) {
        if (tm->used == tm->start.size()) {
            hipEvent_t a, b;
            hipError_t e = hipEventCreate(&a);
            if (e != hipSuccess) return e;
            e = hipEventCreate(&b);
            if (e != hipSuccess) return e;
            tm->start.push_back(a);
            tm->stop.push_back(b);
        }
        hipError_t e = hipEventRecord(tm->start[tm->used], g_state.stream);
        if (e != hipSuccess) return e;
    }
    hipError_t e = launch();
    if (e != hipSuccess) return e;
    if (tm) {
        e = hipEventRecord(tm->stop[tm->used], g_state.stream);
        if (e != hipSuccess) return e;
        tm->used++;
    }
    return hipSuccess;
}

const StencilTiling *class_tiling(int cls)
{
    return g_state.tiling[cls].tz ? &g_state.tiling[cls] : nullptr;
}

template <typename T>
hipError_t stencil(int radius, const T *P, T *PP, const float *V, const Coefs &cf, const Grid &g,
                   const ImageFuse *fuse, const TaperFuse *taper = nullptr)
{
    // kernel class: 0 plain step, 1 fused receiver step, 2 tapered HBC step
    const int cls = fuse ? 1 : (taper ? 2 : 0);
    return timed(cls, [&] {
        return launch_stencil(radius, P, PP, V, cf, g, fuse, taper, g_state.flag,
                              g_state.stream, class_tiling(cls));
    });
}


int check_dispersion()
{
    int h = 0;
    HIPCHK(hipMemcpyAsync(&h, g_state.flag, sizeof h, hipMemcpyDeviceToHost, g_state.stream),
           "reading the dispersion flag");
    HIPCHK(hipStreamSynchronize(g_state.stream), "hipStreamSynchronize");
    if (h) {
        HIPCHK(hipMemsetAsync(g_state.flag, 0, sizeof(int), g_state.stream), "clearing the flag");
        HIPCHK(hipStreamSynchronize(g_state.stream), "hipStreamSynchronize");
        return fail(RTMHIP_ERR_DISPERSION,
                    "RTM ERROR: Numerical Dispersion (NaN produced by the stencil update)");
    }
    return RTMHIP_OK;
}

struct ShotCtx {
    Grid g;
    int radius, hf;
    Coefs cf;
    long long src_off;
    int rx0, ry0, rz;
};

int prepare_shot(const rtmhip_shot *s, const float *coefs, ShotCtx *c)
{
    if (!s || !coefs) return fail(RTMHIP_ERR_INVALID, "shot or coefs is NULL");
    int rc = check_grid(s->nxe, s->nye, s->nze);
    if (rc) return rc;
    rc = check_order(s->order, &c->radius);
    if (rc) return rc;
    c->hf = s->order / 2;
    if (s->blen < c->hf || 2 * s->blen >= s->nxe || 2 * s->blen >= s->nye || 2 * s->blen >= s->nze)
        return fail(RTMHIP_ERR_INVALID, "blen %d must be >= order/2 and leave an interior",
                    s->blen);
    if (s->nt <= 0 || s->ntstep <= 0 || s->nt % s->ntstep != 0)
        return fail(RTMHIP_ERR_INVALID, "ntstep (%d) must divide nt (%d)", s->ntstep, s->nt);
    if (s->sxe < 0 || s->sxe >= s->nxe || s->sye < 0 || s->sye >= s->nye || s->sze < 0 ||
        s->sze >= s->nze)
        return fail(RTMHIP_ERR_INVALID, "source (%d,%d,%d) outside the grid", s->sxe, s->sye,
                    s->sze);
    rc = check_rcv(&s->rcv, s->nze, s->blen);
    if (rc) return rc;
    c->g = dense_grid(s->nxe, s->nye, s->nze);
    c->cf = make_coefs(coefs, s->order);
    c->src_off = ((long long)s->sxe * s->nye + s->sye) * s->nze + s->sze;
    c->rx0 = s->rcv.start_x + s->blen;
    c->ry0 = s->rcv.start_y + s->blen;
    c->rz = s->rcv.depth_z + s->blen;
    return RTMHIP_OK;
}

// The source cell and every receiver cell lie at least blen from each face.
bool off_borders(const rtmhip_shot *shot, const ShotCtx &c)
{
    const int b = shot->blen;
    auto in = [&](long long v, int n) { return v >= b && v < n - b; };
    if (!in(shot->sxe, c.g.nx) || !in(shot->sye, c.g.ny) || !in(shot->sze, c.g.nz)) return false;
    const rtmhip_rcv_geom &r = shot->rcv;
    if (r.count_x <= 0 || r.count_y <= 0) return true;
    return in(c.rz, c.g.nz) && in(c.rx0, c.g.nx) && in(c.ry0, c.g.ny) &&
           in(c.rx0 + (long long)(r.count_x - 1) * r.dist_x, c.g.nx) &&
           in(c.ry0 + (long long)(r.count_y - 1) * r.dist_y, c.g.ny);
}

int check_format(const rtmhip_shot *s)
{
    if (s->wave_format != RTMHIP_WAVE_F32 && s->wave_format != RTMHIP_WAVE_F16 &&
        s->wave_format != RTMHIP_WAVE_BF16)
        return fail(RTMHIP_ERR_INVALID, "unknown wave_format %d", s->wave_format);
    if (s->wave_format != RTMHIP_WAVE_F32 && s->order != 8)
        return fail(RTMHIP_ERR_UNSUPPORTED, "16-bit wavefields need stencil order 8 (got %d)",
                    s->order);
    return RTMHIP_OK;
}

// Calls f(T{}) with T the storage type of the shot's wavefields.
template <typename F>
int with_format(const rtmhip_shot *s, F &&f)
{
    switch (s->wave_format) {
    case RTMHIP_WAVE_F16: return f(_Float16{});
    case RTMHIP_WAVE_BF16: return f(__bf16{});
    default: return f(float{});
    }
}

template <typename T>
T **as(float **p)
{
    return reinterpret_cast<T **>(p);
}

int check_chunk(const rtmhip_shot *s, int lo, int hi)
{
    if (lo < 0 || hi >= s->nt || lo > hi)
        return fail(RTMHIP_ERR_INVALID, "time range [%d,%d] outside [0,%d)", lo, hi, s->nt);
    if (lo / s->ntstep != hi / s->ntstep)
        return fail(RTMHIP_ERR_INVALID, "time range [%d,%d] crosses a receiver chunk of %d", lo,
                    hi, s->ntstep);
    return RTMHIP_OK;
}

}  // namespace

extern "C" {

int rtmhip_version(void) { return 10000; }

const char *rtmhip_last_error(void) { return g_err.c_str(); }

int rtmhip_device_count(int *count)
{
    if (!count) return fail(RTMHIP_ERR_INVALID, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return RTMHIP_OK;
}

int rtmhip_init(int device)
{
    if (g_state.init) {
        if (device == g_state.device) return RTMHIP_OK;
        return fail(RTMHIP_ERR_INVALID, "already initialised on device %d", g_state.device);
    }
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n), "hipGetDeviceCount");
    if (device < 0 || device >= n)
        return fail(RTMHIP_ERR_INVALID, "device %d not present (%d devices)", device, n);
    HIPCHK(hipSetDevice(device), "hipSetDevice");
    if (device < kMaxDevices && !g_scratch[device]) {
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        g_scratch[device] = (BatchShot *)p;
    }
    HIPCHK(hipStreamCreateWithFlags(&g_state.stream, hipStreamNonBlocking), "hipStreamCreate");
    HIPCHK(hipMalloc((void **)&g_state.flag, sizeof(int)), "hipMalloc(flag)");
    HIPCHK(hipMemsetAsync(g_state.flag, 0, sizeof(int), g_state.stream), "hipMemset(flag)");
    HIPCHK(hipStreamSynchronize(g_state.stream), "hipStreamSynchronize");
    g_state.device = device;
    g_state.init = true;
    return RTMHIP_OK;
}

int rtmhip_finalize(void)
{
    if (!g_state.init) return RTMHIP_OK;
    (void)hipStreamSynchronize(g_state.stream);
    for (auto &t : g_state.tm) {
        for (auto e : t.start) (void)hipEventDestroy(e);
        for (auto e : t.stop) (void)hipEventDestroy(e);
        t = Timing();
    }
    (void)hipFree(g_state.flag);
    (void)hipStreamDestroy(g_state.stream);
    g_state = State();
    return RTMHIP_OK;
}

int rtmhip_synchronize(void)
{
    NEED_INIT();
    HIPCHK(hipStreamSynchronize(g_state.stream), "hipStreamSynchronize");
    return check_dispersion();
}

int rtmhip_alloc(void **dptr, size_t bytes)
{
    NEED_INIT();
    if (!dptr) return fail(RTMHIP_ERR_INVALID, "dptr is NULL");
    const hipError_t e = hipMalloc(dptr, bytes ? bytes : 4);
    if (e != hipSuccess) {
        // reported here; cleared so that the next launch check does not see
        // it (callers may fall back to a smaller allocation)
        (void)hipGetLastError();
        *dptr = nullptr;
        return hip_fail(e, "hipMalloc");
    }
    return RTMHIP_OK;
}

int rtmhip_free(void *dptr)
{
    NEED_INIT();
    if (!dptr) return RTMHIP_OK;
    HIPCHK(hipStreamSynchronize(g_state.stream), "hipStreamSynchronize");
    HIPCHK(hipFree(dptr), "hipFree");
    return RTMHIP_OK;
}

int rtmhip_copy_to_device(void *dst, const void *src, size_t bytes)
{
    NEED_INIT();
    if (!dst || !src) return fail(RTMHIP_ERR_INVALID, "NULL pointer");
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g_state.stream), "copy H2D");
    HIPCHK(hipStreamSynchronize(g_state.stream), "hipStreamSynchronize");
    return RTMHIP_OK;
}

int rtmhip_copy_from_device(void *dst, const void *src, size_t bytes)
{
    NEED_INIT();
    if (!dst || !src) return fail(RTMHIP_ERR_INVALID, "NULL pointer");
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g_state.stream), "copy D2H");
    HIPCHK(hipStreamSynchronize(g_state.stream), "hipStreamSynchronize");
    return RTMHIP_OK;
}

int rtmhip_copy_device(void *dst, const void *src, size_t bytes)
{
    NEED_INIT();
    if (!dst || !src) return fail(RTMHIP_ERR_INVALID, "NULL pointer");
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, g_state.stream), "copy D2D");
    return RTMHIP_OK;
}

int rtmhip_fill(float *dptr, size_t count, float value)
{
    NEED_INIT();
    if (!dptr && count) return fail(RTMHIP_ERR_INVALID, "NULL pointer");
    HIPCHK(launch_fill(dptr, count, value, g_state.stream), "fill");
    return RTMHIP_OK;
}

int rtmhip_fill16(void *dptr, size_t count, unsigned short bits)
{
    NEED_INIT();
    if (!dptr && count) return fail(RTMHIP_ERR_INVALID, "NULL pointer");
    HIPCHK(launch_fill_u16(static_cast<uint16_t *>(dptr), count, bits, g_state.stream), "fill16");
    return RTMHIP_OK;
}

int rtmhip_step(const float *P, float *PP, const float *V2DT2, const float *coefs, int nxe,
                int nye, int nze, int order)
{
    NEED_INIT();
    int radius, rc;
    if ((rc = check_order(order, &radius))) return rc;
    if ((rc = check_grid(nxe, nye, nze))) return rc;
    if (!P || !PP || !V2DT2 || !coefs) return fail(RTMHIP_ERR_INVALID, "NULL pointer");
    if (P == PP) return fail(RTMHIP_ERR_INVALID, "P and PP must be distinct buffers");
    const Coefs cf = make_coefs(coefs, order);
    HIPCHK(stencil(radius, P, PP, V2DT2, cf, dense_grid(nxe, nye, nze), nullptr), "stencil");
    return RTMHIP_OK;
}

int rtmhip_taper(float *P, const float *taper, int nxe, int nye, int nze, int blen,
                 int upper_only)
{
    NEED_INIT();
    int rc;
    if ((rc = check_grid(nxe, nye, nze))) return rc;
    if (!P || (!taper && blen > 0)) return fail(RTMHIP_ERR_INVALID, "NULL pointer");
    if (blen < 0 || 2 * blen >= nxe || 2 * blen >= nye || 2 * blen >= nze)
        return fail(RTMHIP_ERR_INVALID, "blen %d leaves no interior", blen);
    HIPCHK(launch_taper(P, taper, dense_grid(nxe, nye, nze), blen, upper_only != 0,
                        g_state.stream),
           "taper");
    return RTMHIP_OK;
}

int rtmhip_apply_source(float *PP, int nxe, int nye, int nze, int sxe, int sye, int sze,
                        float value)
{
    NEED_INIT();
    int rc;
    if ((rc = check_grid(nxe, nye, nze))) return rc;
    if (!PP) return fail(RTMHIP_ERR_INVALID, "NULL pointer");
    if (sxe < 0 || sxe >= nxe || sye < 0 || sye >= nye || sze < 0 || sze >= nze)
        return RTMHIP_OK;  // outside the grid: no-op, as getWithGlobalCoordinate
    const long long o = ((long long)sxe * nye + sye) * nze + sze;
    HIPCHK(launch_source(PP, o, nullptr, 0, value, g_state.stream), "source");
    return RTMHIP_OK;
}

int rtmhip_receivers(float *PP, float *rcv, int nxe, int nye, int nze, int blen,
                     const rtmhip_rcv_geom *geom, int ntstep, int lt, int restore)
{
    NEED_INIT();
    int rc;
    if ((rc = check_grid(nxe, nye, nze))) return rc;
    if ((rc = check_rcv(geom, nze, blen))) return rc;
    if (!PP || !rcv) return fail(RTMHIP_ERR_INVALID, "NULL pointer");
    if (ntstep <= 0 || lt < 0 || lt >= ntstep)
        return fail(RTMHIP_ERR_INVALID, "lt %d outside [0,%d)", lt, ntstep);
    HIPCHK(launch_receivers(PP, rcv, dense_grid(nxe, nye, nze), geom->start_x + blen,
                            geom->start_y + blen, geom->depth_z + blen, geom->dist_x,
                            geom->dist_y, geom->count_x, geom->count_y, ntstep, lt,
                            restore != 0, g_state.stream),
           "receivers");
    return RTMHIP_OK;
}

int rtmhip_upper_border(float *PP, float *upb, int nxe, int nye, int nze, int blen, int order,
                        int lt, int restore)
{
    NEED_INIT();
    int radius, rc;
    if ((rc = check_grid(nxe, nye, nze))) return rc;
    if ((rc = check_order(order, &radius))) return rc;
    if (!PP || !upb) return fail(RTMHIP_ERR_INVALID, "NULL pointer");
    if (blen < radius || blen > nze) return fail(RTMHIP_ERR_INVALID, "blen %d < order/2", blen);
    if (lt < 0) return fail(RTMHIP_ERR_INVALID, "negative lt");
    HIPCHK(launch_upb(PP, upb, dense_grid(nxe, nye, nze), blen, radius, lt, restore != 0,
                      g_state.stream),
           "upper border");
    return RTMHIP_OK;
}

int rtmhip_image_condition(float *img, const float *PS, const float *PR, int nxe, int nye,
                           int nze, int blen)
{
    NEED_INIT();
    int rc;
    if ((rc = check_grid(nxe, nye, nze))) return rc;
    if (!img || !PS || !PR) return fail(RTMHIP_ERR_INVALID, "NULL pointer");
    if (blen < 0) return fail(RTMHIP_ERR_INVALID, "negative blen");
    HIPCHK(launch_image(img, PS, PR, dense_grid(nxe, nye, nze), blen, nze - blen, g_state.stream),
           "image");
    return RTMHIP_OK;
}

int rtmhip_freq_image(float *img, const float *PS, const float *PR, float *sre, float *sim,
                      float *rre, float *rim, const float *kre, const float *kim,
                      const float *w2, int nxe, int nye, int nze, int blen, int nwstep, int iw0,
                      int nw, int nt, int st)
{
    NEED_INIT();
    int rc;
    if ((rc = check_grid(nxe, nye, nze))) return rc;
    if (!img || !PS || !PR || !sre || !sim || !rre || !rim || !kre || !kim || !w2)
        return fail(RTMHIP_ERR_INVALID, "NULL pointer");
    if (blen < 0 || 2 * blen >= nxe || 2 * blen >= nye || 2 * blen >= nze)
        return fail(RTMHIP_ERR_INVALID, "blen %d leaves no interior", blen);
    if (nwstep < 1 || nw < 0 || iw0 < 0 || nt <= 0 || st < 0 || st >= nt)
        return fail(RTMHIP_ERR_INVALID, "bad frequency group or time index");
    HIPCHK(launch_freq_image(img, PS, PR, sre, sim, rre, rim, kre, kim, w2,
                             dense_grid(nxe, nye, nze), blen, nwstep, iw0, nw, nt, st,
                             g_state.stream),
           "frequency contributions");
    return RTMHIP_OK;
}

int rtmhip_freq_steps(const rtmhip_shot *shot, const float *V2DT2, const float *coefs,
                      const float *taper, const float *ricker, float **Ps, float **PPs,
                      float **Pr, float **PPr, const float *rcv_chunk, float *img, float *sre,
                      float *sim, float *rre, float *rim, const float *kre, const float *kim,
                      const float *w2, int nwstep, int iw0, int nw, int it_hi, int nsteps)
{
    NEED_INIT();
    ShotCtx c;
    int rc;
    if ((rc = prepare_shot(shot, coefs, &c))) return rc;
    if (shot->wave_format != RTMHIP_WAVE_F32)
        return fail(RTMHIP_ERR_UNSUPPORTED, "frequency-domain steps take FP32 wavefields");
    if (!V2DT2 || !taper || !ricker || !Ps || !PPs || !Pr || !PPr || !*Ps || !*PPs || !*Pr ||
        !*PPr || !rcv_chunk || !img || !sre || !sim || !rre || !rim || !kre || !kim || !w2)
        return fail(RTMHIP_ERR_INVALID, "NULL pointer");
    const int nt = shot->nt, b = shot->blen;
    if (nsteps < 0 || it_hi >= nt || it_hi - nsteps + 1 < 0)
        return fail(RTMHIP_ERR_INVALID, "steps %d..%d outside [0, %d)", it_hi - nsteps + 1, it_hi,
                    nt);
    if (nwstep < 1 || nw < 0 || iw0 < 0)
        return fail(RTMHIP_ERR_INVALID, "bad frequency group");
    if (nsteps > 0 && it_hi / shot->ntstep != (it_hi - nsteps + 1) / shot->ntstep)
        return fail(RTMHIP_ERR_INVALID, "steps %d..%d span two receiver chunks of %d",
                    it_hi - nsteps + 1, it_hi, shot->ntstep);
    hipStream_t s = g_state.stream;
    // taper3D of the four grids after the steps, as the reference orders it,
    // or applied to the new fields' output (the modeling loop's output form,
    // when the every-border taper takes it for FP32 and the source and the
    // receivers lie off the borders): each new field is stored tapered (with
    // the x-y frame and the front-right edge by their kernels), the older
    // field, which the reference tapers in place after the step, is tapered
    // at the next step's centre load instead (not on the call's first step,
    // whose older fields memory already holds tapered), and in place after
    // the call's last step.
    const bool out = c.radius == 4 && b > 0 && g_state.taper_all_mode[RTMHIP_WAVE_F32] == 2 &&
                     off_borders(shot, c);
    for (int k = 0; k < nsteps; ++k) {
        const int rt = it_hi - k, st = nt - 1 - rt, lt = rt % shot->ntstep;
        if (out) {
            const TaperFuse tf{taper, b, k == 0 ? 0 : 1, 2};
            float **fields[2][2] = {{Ps, PPs}, {Pr, PPr}};
            for (auto &f : fields) {
                HIPCHK(launch_taper_frame_xy(*f[1], taper, c.g, b, c.radius, k == 0 ? 1 : 2, s),
                       "frame taper");
                HIPCHK(stencil(c.radius, (const float *)*f[0], *f[1], V2DT2, c.cf, c.g, nullptr,
                               &tf),
                       "stencil");
                HIPCHK(launch_taper_quirk_out(*f[1], taper, c.g, b, s), "front-right taper");
            }
        } else {
            HIPCHK(stencil(c.radius, (const float *)*Ps, *PPs, V2DT2, c.cf, c.g, nullptr),
                   "source stencil");
            HIPCHK(stencil(c.radius, (const float *)*Pr, *PPr, V2DT2, c.cf, c.g, nullptr),
                   "receiver stencil");
            HIPCHK(launch_taper_all2(*Ps, *PPs, taper, c.g, b, s), "taper");
            HIPCHK(launch_taper_all2(*Pr, *PPr, taper, c.g, b, s), "taper");
        }
        HIPCHK(launch_source(*PPs, c.src_off, ricker, st, 0.f, s), "source");
        HIPCHK(launch_receivers(*PPr, const_cast<float *>(rcv_chunk), c.g, c.rx0, c.ry0, c.rz,
                                shot->rcv.dist_x, shot->rcv.dist_y, shot->rcv.count_x,
                                shot->rcv.count_y, shot->ntstep, lt, true, s),
               "receivers");
        HIPCHK(launch_freq_image(img, *PPs, *PPr, sre, sim, rre, rim, kre, kim, w2, c.g, b, nwstep,
                                 iw0, nw, nt, st, s),
               "frequency contributions");
        std::swap(*Ps, *PPs);
        std::swap(*Pr, *PPr);
    }
    if (out && nsteps > 0) HIPCHK(launch_taper_all2(*PPs, *PPr, taper, c.g, b, s), "taper");
    return check_dispersion();
}

int rtmhip_image_finish(const float *img, float *out, int nxe, int nye, int nze, int blen,
                        const float *fcoefs)
{
    NEED_INIT();
    int rc;
    if ((rc = check_grid(nxe, nye, nze))) return rc;
    if (!img || !out || !fcoefs) return fail(RTMHIP_ERR_INVALID, "NULL pointer");
    if (blen < 1 || 2 * blen >= nxe || 2 * blen >= nye || 2 * blen >= nze)
        return fail(RTMHIP_ERR_INVALID, "blen %d must be >= 1 and leave an interior", blen);
    const Coefs fc = make_coefs(fcoefs, 2);
    HIPCHK(launch_image_finish(img, out, dense_grid(nxe, nye, nze), blen, fc, g_state.stream),
           "image finish");
    return RTMHIP_OK;
}

int rtmhip_stack(float *dst, const float *src, size_t count)
{
    NEED_INIT();
    if ((!dst || !src) && count) return fail(RTMHIP_ERR_INVALID, "NULL pointer");
    HIPCHK(launch_stack(dst, src, count, g_state.stream), "stack");
    return RTMHIP_OK;
}

}  // extern "C"

namespace {


template <typename T>
int model_loop(const rtmhip_shot *shot, const ShotCtx &c, const float *V2DT2, const float *taper,
               const float *ricker, T **P, T **PP, float *rcv_chunk, int it0, int nsteps)
{
    hipStream_t s = g_state.stream;
    // taper3D of both fields before each step (order 8) runs in one of three
    // forms, with identical results.  By default (and with
    // rtmhip_set_taper_fusion mode 3) it is applied to each step's output:
    // the step stores taper3D of its new value (the next step's P, so P is
    // read as stored) and tapers the centre PP once; PP's x-y frame is
    // tapered twice before the step (it is this step's PP and the next
    // step's P), the z frame by the stencil, and the front-right edge is
    // written after the step from the stored (b, y, z).  The first step of a
    // call tapers P in place, the last tapers PP in place and runs the plain
    // step, so the state between calls is the reference's.  The source and
    // the receivers must lie off the borders, where taper3D leaves values
    // unchanged, since they act on the stored (tapered) field; otherwise the
    // separate pass runs.  Salt grid, 300 steps: FP32 0.140 s against 0.187 s
    // for the separate pass, FP16 0.120 against 0.128.
    // Mode 2 fuses it on load instead (P tapered on load, the centre PP
    // `pp_taps` times, PP's frame tapered in place before the step, the taper
    // still pending on P applied after the last step): a taper chain on every
    // loaded value, 0.183 s.  Modes 0 and 1 run the separate in-place pass.
    const int b = shot->blen;
    const int amode = (c.radius == 4 && b > 0) ? g_state.taper_all_mode[shot->wave_format] : 0;
    const bool fused = amode == 1;
    const bool out = amode == 2 && off_borders(shot, c);
    for (int it = it0; it < it0 + nsteps; ++it) {
        T *t = *P;
        *P = *PP;
        *PP = t;
        const bool last = it == it0 + nsteps - 1;
        if (out && !last) {
            if (it == it0) HIPCHK(launch_taper(*P, taper, c.g, b, false, s), "taper");
            const TaperFuse tf{taper, b, 1, 2};
            HIPCHK(launch_taper_frame_xy(*PP, taper, c.g, b, c.radius, 2, s), "frame taper");
            HIPCHK(stencil(c.radius, *P, *PP, V2DT2, c.cf, c.g, nullptr, &tf), "stencil");
            HIPCHK(launch_taper_quirk_out(*PP, taper, c.g, b, s), "front-right taper");
        } else if (out) {
            if (it == it0)
                HIPCHK(launch_taper_all2(*P, *PP, taper, c.g, b, s), "taper");
            else
                HIPCHK(launch_taper(*PP, taper, c.g, b, false, s), "taper");
            HIPCHK(stencil(c.radius, *P, *PP, V2DT2, c.cf, c.g, nullptr), "stencil");
        } else if (fused) {
            const TaperFuse tf{taper, b, it == it0 ? 1 : 2, 1};
            HIPCHK(launch_taper_frame_all(*PP, taper, c.g, b, c.radius, tf.pp_taps, s),
                   "frame taper");
            HIPCHK(stencil(c.radius, *P, *PP, V2DT2, c.cf, c.g, nullptr, &tf), "stencil");
        } else {
            // taperAllBorders of P, then of PP: disjoint cells, one pass
            HIPCHK(launch_taper_all2(*P, *PP, taper, c.g, b, s), "taper");
            HIPCHK(stencil(c.radius, *P, *PP, V2DT2, c.cf, c.g, nullptr), "stencil");
        }
        HIPCHK(launch_source(*PP, c.src_off, ricker, it, 0.f, s), "source");
        HIPCHK(launch_receivers(*PP, rcv_chunk, c.g, c.rx0, c.ry0, c.rz, shot->rcv.dist_x,
                                shot->rcv.dist_y, shot->rcv.count_x, shot->rcv.count_y,
                                shot->ntstep, it % shot->ntstep, false, s),
               "receivers");
    }
    if (fused && nsteps > 0) HIPCHK(launch_taper(*P, taper, c.g, b, false, s), "taper");
    return RTMHIP_OK;
}

template <typename T>
int forward_loop(const rtmhip_shot *shot, const ShotCtx &c, int hbc, const float *V2DT2,
                 const float *taper, const float *ricker, T **P, T **PP, T *upb, int it0,
                 int nsteps)
{
    hipStream_t s = g_state.stream;
    // HBC, three forms of the two in-place upper tapers of each step:
    // * on the output (the F32 default, rtmhip_set_taper_fusion mode 4): each
    //   step stores the upper taper of its new value, the next step's P, so
    //   P is read as stored and only the centre PP is tapered on load; the
    //   planes [blen-hf, blen) are stored untapered, the upper-border history
    //   is saved from them and they are tapered in place after the source;
    //   PP's frame is tapered twice (once in those planes).  The first step
    //   of a call tapers P in place, the last tapers PP in place and runs the
    //   plain step, so the state between calls is the reference's.  Needs the
    //   source at z >= blen (it acts on the stored field).
    // * fused on load (mode 1): P tapered on load, the centre PP tapered
    //   twice -- once on the first step of a call, whose PP already carries
    //   one taper; the taper still pending on P is applied in place after the
    //   last step, so the state between calls is the reference's.
    // * separate (mode 0, the 16-bit default): the two tapers run in place
    //   (one pass over both fields) before a plain step, as the reference
    //   orders them.
    const bool sep = g_state.taper_sep[shot->wave_format];
    const int b = shot->blen;
    const bool outu = hbc && g_state.taper_outu[shot->wave_format] && c.radius == 4 &&
                      shot->sze >= b;
    for (int it = it0; it < it0 + nsteps; ++it) {
        // history layer: every step held (hbc 1) or one receiver chunk (hbc 2)
        const int ul = hbc == 2 ? it % shot->ntstep : it;
        T *t = *P;
        *P = *PP;
        *PP = t;
        if (outu) {
            if (it < it0 + nsteps - 1) {
                if (it == it0) HIPCHK(launch_taper(*P, taper, c.g, b, true, s), "taper");
                const TaperFuse tf{taper, b, 1, 3, c.hf};
                HIPCHK(stencil(c.radius, *P, *PP, V2DT2, c.cf, c.g, nullptr, &tf), "stencil");
                HIPCHK(launch_taper_frame(*PP, taper, c.g, b, c.radius, 2, s, c.hf),
                       "frame taper");
                HIPCHK(launch_source(*PP, c.src_off, ricker, it, 0.f, s), "source");
                HIPCHK(launch_upb_save_taper(*PP, upb, taper, c.g, b, c.hf, ul, s), "upb save");
                continue;
            }
            if (it == it0)
                HIPCHK(launch_taper_upper2(*P, *PP, taper, c.g, b, s), "taper");
            else
                HIPCHK(launch_taper(*PP, taper, c.g, b, true, s), "taper");
            HIPCHK(stencil(c.radius, *P, *PP, V2DT2, c.cf, c.g, nullptr), "stencil");
        } else if (hbc && sep) {
            HIPCHK(launch_taper_upper2(*P, *PP, taper, c.g, shot->blen, s), "taper");
            HIPCHK(stencil(c.radius, *P, *PP, V2DT2, c.cf, c.g, nullptr), "stencil");
        } else if (hbc) {
            const TaperFuse tf{taper, shot->blen, it == it0 ? 1 : 2};
            HIPCHK(stencil(c.radius, *P, *PP, V2DT2, c.cf, c.g, nullptr, &tf), "stencil");
            HIPCHK(launch_taper_frame(*PP, taper, c.g, shot->blen, c.radius, tf.pp_taps, s),
                   "frame taper");
        } else {
            HIPCHK(stencil(c.radius, *P, *PP, V2DT2, c.cf, c.g, nullptr), "stencil");
        }
        HIPCHK(launch_source(*PP, c.src_off, ricker, it, 0.f, s), "source");
        if (hbc) HIPCHK(launch_upb(*PP, upb, c.g, shot->blen, c.hf, ul, false, s), "upb save");
    }
    if (hbc && !sep && !outu) HIPCHK(launch_taper(*P, taper, c.g, shot->blen, true, s), "taper");
    return RTMHIP_OK;
}

template <typename T>
int backward_loop(const rtmhip_shot *shot, const ShotCtx &c, int hbc, const float *V2DT2,
                  const float *taper, T **Ps, T **PPs, T **Pr, T **PPr, const T *upb,
                  const float *rcv_chunk, float *img, int it_hi, int nsteps, int fused)
{
    hipStream_t s = g_state.stream;
    const int nt = shot->nt, b = shot->blen;
    const int z0 = b, z1 = shot->nze - b;
    const bool dual = fused == 2 && std::is_same<T, float>::value &&
                      dual_step_supported(c.radius, shot->wave_format);
    // separate in-place receiver tapers before an untapered fused step (not
    // with the dual kernel, which fuses them)
    const bool sep = hbc && fused && !dual && g_state.taper_sep[shot->wave_format];
    for (int it = it_hi; it > it_hi - nsteps; --it) {
        const int lt = it % shot->ntstep;
        const int ul = hbc == 2 ? lt : it;  // history layer, as in forward_loop
        if constexpr (std::is_same<T, float>::value) {
            if (dual && it < nt - 2) {
                // both backward steps in one pass (class 3), then the UPB
                // restore of the source field, its swap, and the receiver
                // field's frame taper, injection and imaging as below
                const TaperFuse tf{taper, b, it == it_hi ? 1 : 2};
                ImageFuse fz{};
                fz.img = img;
                fz.z0 = z0;
                fz.z1 = z1;
                fz.rz = c.rz;
                fz.rx0 = c.rx0;
                fz.rdx = std::max(1, shot->rcv.dist_x);
                fz.rnx = shot->rcv.count_x;
                fz.ry0 = c.ry0;
                fz.rdy = std::max(1, shot->rcv.dist_y);
                fz.rny = shot->rcv.count_y;
                HIPCHK(timed(3, [&] {
                           return launch_stencil_dual(c.radius, *Ps, *PPs, *Pr, *PPr, V2DT2, c.cf,
                                                      c.g, fz, hbc ? &tf : nullptr, g_state.flag,
                                                      s, class_tiling(3));
                       }),
                       "dual stencil");
                if (hbc)
                    HIPCHK(launch_upb(*PPs, const_cast<T *>(upb), c.g, b, c.hf, ul, true, s),
                           "upb restore");
                T *t = *Ps;
                *Ps = *PPs;
                *PPs = t;
                if (hbc)
                    HIPCHK(launch_taper_frame(*PPr, taper, c.g, b, c.radius, tf.pp_taps, s),
                           "frame taper");
                HIPCHK(launch_receivers_image(*PPr, rcv_chunk, c.g, c.rx0, c.ry0, c.rz,
                                              shot->rcv.dist_x, shot->rcv.dist_y,
                                              shot->rcv.count_x, shot->rcv.count_y,
                                              shot->ntstep, lt, img, (const T *)*Ps, z0, z1,
                                              c.radius, s),
                       "receivers+image");
                t = *Pr;
                *Pr = *PPr;
                *PPr = t;
                continue;
            }
        }
        // Source wavefield at time it (RTMFiniteDifferencesKernel_HBC.cpp:283-309).
        const T *S;
        if (it == nt - 1) {
            S = *PPs;  // snap1
        } else if (it == nt - 2) {
            S = *Ps;   // snap0 (entry state: Ps = snap0, PPs = snap1)
        } else {
            HIPCHK(stencil(c.radius, *Ps, *PPs, V2DT2, c.cf, c.g, nullptr), "stencil");
            if (hbc)
                HIPCHK(launch_upb(*PPs, const_cast<T *>(upb), c.g, b, c.hf, ul, true, s),
                       "upb restore");
            T *t = *Ps;
            *Ps = *PPs;
            *PPs = t;
            S = *Ps;
        }
        // Receiver wavefield and imaging (:311-345).
        if (hbc && !fused) HIPCHK(launch_taper_upper2(*Pr, *PPr, taper, c.g, b, s), "taper");
        if (fused) {
            // tapers fused as in the forward loop (finishing taper below), or
            // applied in place first
            if (sep) HIPCHK(launch_taper_upper2(*Pr, *PPr, taper, c.g, b, s), "taper");
            const TaperFuse tf{taper, b, it == it_hi ? 1 : 2};
            ImageFuse fz;
            fz.img = img;
            fz.ps = S;
            fz.z0 = z0;
            fz.z1 = z1;
            fz.rz = c.rz;
            fz.rx0 = c.rx0;
            fz.rdx = std::max(1, shot->rcv.dist_x);
            fz.rnx = shot->rcv.count_x;
            fz.ry0 = c.ry0;
            fz.rdy = std::max(1, shot->rcv.dist_y);
            fz.rny = shot->rcv.count_y;
            HIPCHK(stencil(c.radius, *Pr, *PPr, V2DT2, c.cf, c.g, &fz,
                           hbc && !sep ? &tf : nullptr),
                   "fused stencil");
            if (hbc && !sep)
                HIPCHK(launch_taper_frame(*PPr, taper, c.g, b, c.radius, tf.pp_taps, s),
                       "frame taper");
            HIPCHK(launch_receivers_image(*PPr, rcv_chunk, c.g, c.rx0, c.ry0, c.rz,
                                          shot->rcv.dist_x, shot->rcv.dist_y, shot->rcv.count_x,
                                          shot->rcv.count_y, shot->ntstep, lt, img, S, z0, z1,
                                          c.radius, s),
                   "receivers+image");
        } else {
            HIPCHK(stencil(c.radius, *Pr, *PPr, V2DT2, c.cf, c.g, nullptr), "stencil");
            HIPCHK(launch_receivers(*PPr, const_cast<float *>(rcv_chunk), c.g, c.rx0, c.ry0,
                                    c.rz, shot->rcv.dist_x, shot->rcv.dist_y, shot->rcv.count_x,
                                    shot->rcv.count_y, shot->ntstep, lt, true, s),
                   "receivers");
            HIPCHK(launch_image(img, S, (const T *)*PPr, c.g, z0, z1, s), "image");
        }
        T *t = *Pr;
        *Pr = *PPr;
        *PPr = t;
    }
    // pending in-place taper of the older receiver field (fused HBC path)
    if (hbc && fused && !sep) HIPCHK(launch_taper(*PPr, taper, c.g, b, true, s), "taper");
    return RTMHIP_OK;
}

// Snapshot slot of time step it, or -1: the snapshot steps are nt-1,
// nt-1-k, nt-1-2k, ... (slot 0 is the last step, imaged first).
inline int snap_slot(int it, int nt, int k)
{
    const int d = nt - 1 - it;
    return d % k == 0 ? d / k : -1;
}

inline size_t snap_cells(const rtmhip_shot *shot)
{
    return (size_t)shot->nxe * shot->nye * (size_t)(shot->nze - 2 * shot->blen);
}

// Forward propagation of snapshot RTM: the HBC forward step sequence of
// forward_loop (upper tapers fused or separate, frame taper, source) with
// the source wavefield of every snapshot step handed to store(PP, slot)
// instead of the upper-border history.
template <typename T, typename Store>
int snap_forward_run(const rtmhip_shot *shot, const ShotCtx &c, const float *V2DT2,
                     const float *taper, const float *ricker, T **P, T **PP, int k, int it0,
                     int nsteps, Store &&store)
{
    hipStream_t s = g_state.stream;
    const int b = shot->blen;
    const bool sep = g_state.taper_sep[shot->wave_format];
    // the upper taper on the output as in forward_loop (no history planes
    // held: the snapshots cover z >= blen only)
    const bool outu = g_state.taper_outu[shot->wave_format] && c.radius == 4 && shot->sze >= b;
    for (int it = it0; it < it0 + nsteps; ++it) {
        T *t = *P;
        *P = *PP;
        *PP = t;
        if (outu && it < it0 + nsteps - 1) {
            if (it == it0) HIPCHK(launch_taper(*P, taper, c.g, b, true, s), "taper");
            const TaperFuse tf{taper, b, 1, 3, 0};
            HIPCHK(stencil(c.radius, *P, *PP, V2DT2, c.cf, c.g, nullptr, &tf), "stencil");
            HIPCHK(launch_taper_frame(*PP, taper, c.g, b, c.radius, 2, s), "frame taper");
        } else if (outu) {
            if (it == it0)
                HIPCHK(launch_taper_upper2(*P, *PP, taper, c.g, b, s), "taper");
            else
                HIPCHK(launch_taper(*PP, taper, c.g, b, true, s), "taper");
            HIPCHK(stencil(c.radius, *P, *PP, V2DT2, c.cf, c.g, nullptr), "stencil");
        } else if (sep) {
            HIPCHK(launch_taper_upper2(*P, *PP, taper, c.g, b, s), "taper");
            HIPCHK(stencil(c.radius, *P, *PP, V2DT2, c.cf, c.g, nullptr), "stencil");
        } else {
            const TaperFuse tf{taper, b, it == it0 ? 1 : 2};
            HIPCHK(stencil(c.radius, *P, *PP, V2DT2, c.cf, c.g, nullptr, &tf), "stencil");
            HIPCHK(launch_taper_frame(*PP, taper, c.g, b, c.radius, tf.pp_taps, s),
                   "frame taper");
        }
        HIPCHK(launch_source(*PP, c.src_off, ricker, it, 0.f, s), "source");
        const int j = snap_slot(it, shot->nt, k);
        if (j >= 0) {
            const int rc = store((const T *)*PP, j);
            if (rc) return rc;
        }
    }
    if (!sep && !outu) HIPCHK(launch_taper(*P, taper, c.g, b, true, s), "taper");
    return RTMHIP_OK;
}

template <typename T, typename S>
int snap_forward_loop(const rtmhip_shot *shot, const ShotCtx &c, const float *V2DT2,
                      const float *taper, const float *ricker, T **P, T **PP, S *snaps, int k,
                      int it0, int nsteps)
{
    const size_t per = snap_cells(shot);
    return snap_forward_run(shot, c, V2DT2, taper, ricker, P, PP, k, it0, nsteps,
                            [&](const T *f, int j) {
                                HIPCHK(launch_snap_store(f, snaps + (size_t)j * per, c.g,
                                                         shot->blen, g_state.stream),
                                       "snapshot");
                                return RTMHIP_OK;
                            });
}

// Backward pass of snapshot RTM: only the receiver field propagates (the
// receiver half of backward_loop: upper tapers, step, receiver injection);
// at every snapshot step image(PR, slot) applies img += k * (snapshot * PR).
template <typename T, typename Image>
int snap_backward_run(const rtmhip_shot *shot, const ShotCtx &c, const float *V2DT2,
                      const float *taper, T **Pr, T **PPr, int k, const float *rcv_chunk,
                      int it_hi, int nsteps, Image &&image)
{
    hipStream_t s = g_state.stream;
    const int b = shot->blen;
    const bool sep = g_state.taper_sep[shot->wave_format];
    // the receiver-only steps take the upper taper on the output like the
    // forward loops (the receivers and the imaged range lie at z >= blen)
    const bool outu = g_state.taper_outu[shot->wave_format] && c.radius == 4 && c.rz >= b;
    for (int it = it_hi; it > it_hi - nsteps; --it) {
        if (outu && it > it_hi - nsteps + 1) {
            if (it == it_hi) HIPCHK(launch_taper(*Pr, taper, c.g, b, true, s), "taper");
            const TaperFuse tf{taper, b, 1, 3, 0};
            HIPCHK(stencil(c.radius, *Pr, *PPr, V2DT2, c.cf, c.g, nullptr, &tf), "stencil");
            HIPCHK(launch_taper_frame(*PPr, taper, c.g, b, c.radius, 2, s), "frame taper");
        } else if (outu) {
            if (it == it_hi)
                HIPCHK(launch_taper_upper2(*Pr, *PPr, taper, c.g, b, s), "taper");
            else
                HIPCHK(launch_taper(*PPr, taper, c.g, b, true, s), "taper");
            HIPCHK(stencil(c.radius, *Pr, *PPr, V2DT2, c.cf, c.g, nullptr), "stencil");
        } else if (sep) {
            HIPCHK(launch_taper_upper2(*Pr, *PPr, taper, c.g, b, s), "taper");
            HIPCHK(stencil(c.radius, *Pr, *PPr, V2DT2, c.cf, c.g, nullptr), "stencil");
        } else {
            const TaperFuse tf{taper, b, it == it_hi ? 1 : 2};
            HIPCHK(stencil(c.radius, *Pr, *PPr, V2DT2, c.cf, c.g, nullptr, &tf), "stencil");
            HIPCHK(launch_taper_frame(*PPr, taper, c.g, b, c.radius, tf.pp_taps, s),
                   "frame taper");
        }
        HIPCHK(launch_receivers(*PPr, const_cast<float *>(rcv_chunk), c.g, c.rx0, c.ry0, c.rz,
                                shot->rcv.dist_x, shot->rcv.dist_y, shot->rcv.count_x,
                                shot->rcv.count_y, shot->ntstep, it % shot->ntstep, true, s),
               "receivers");
        const int j = snap_slot(it, shot->nt, k);
        if (j >= 0) {
            const int rc = image((const T *)*PPr, j);
            if (rc) return rc;
        }
        T *t = *Pr;
        *Pr = *PPr;
        *PPr = t;
    }
    if (!sep && !outu) HIPCHK(launch_taper(*PPr, taper, c.g, b, true, s), "taper");
    return RTMHIP_OK;
}

template <typename T, typename S>
int snap_backward_loop(const rtmhip_shot *shot, const ShotCtx &c, const float *V2DT2,
                       const float *taper, T **Pr, T **PPr, const S *snaps, int k,
                       const float *rcv_chunk, float *img, int it_hi, int nsteps)
{
    const size_t per = snap_cells(shot);
    return snap_backward_run(shot, c, V2DT2, taper, Pr, PPr, k, rcv_chunk, it_hi, nsteps,
                             [&](const T *pr, int j) {
                                 HIPCHK(launch_snap_image(img, snaps + (size_t)j * per, pr, c.g,
                                                          shot->blen, (float)k, g_state.stream),
                                        "snapshot image");
                                 return RTMHIP_OK;
                             });
}

// Transform-coded snapshots: one slot per snapshot step, each the entry
// pool (K 32-bit entries), the brick table (3 words per brick) and 3 words
// of metadata, every part 256-byte aligned; the work buffer holds the
// coefficients of one snapshot ([bricks][2048] FP32) and the 256-bin
// exponent histogram.
struct DctSlots {
    DctLayout L;
    long long K;              // coefficient budget per snapshot
    size_t base;              // byte offset of slot 0 (the basis precedes it)
    size_t entries, table, meta, slot;  // byte offsets in a slot, slot size
    size_t coef, hist, work;            // byte offsets in work, work size
};

inline size_t align256(size_t n) { return (n + 255) / 256 * 256; }

int dct_slots(const Grid &g, int blen, int keep_div, DctSlots *d)
{
    if (keep_div < 1) return fail(RTMHIP_ERR_INVALID, "keep_div %d below 1", keep_div);
    d->L = dct_layout(g);
    const long long cells = (long long)g.nx * g.ny * (long long)(g.nz - 2 * blen);
    d->K = std::max(1LL, cells / keep_div);
    if (d->K >= (1LL << 31) || d->L.nbr * 3 >= (1LL << 31))
        return fail(RTMHIP_ERR_INVALID, "snapshot too large for 32-bit entry offsets");
    d->base = align256(sizeof(BrickBasis)) + 4096;
    d->entries = 0;
    d->table = align256((size_t)d->K * 4);
    d->meta = d->table + align256((size_t)d->L.nbr * 12);
    d->slot = d->meta + 256;
    d->coef = 0;
    d->hist = align256((size_t)d->L.nbr * 2048 * 4);
    d->work = d->hist + 1024;
    return RTMHIP_OK;
}

template <typename T>
hipError_t dct_store(const T *f, const Grid &g, int b, const DctSlots &d, void *store, int j,
                     void *work)
{
    char *sl = static_cast<char *>(store) + d.base + (size_t)j * d.slot;
    char *wk = static_cast<char *>(work);
    return launch_dct_encode(f, g, b, g.nz - b, d.L, static_cast<const BrickBasis *>(store), d.K,
                             (float *)(wk + d.coef), (unsigned *)(wk + d.hist),
                             (unsigned *)(sl + d.meta), (unsigned *)(sl + d.entries),
                             (unsigned *)(sl + d.table), g_state.stream);
}

template <typename T>
hipError_t dct_image(const void *store, int j, const DctSlots &d, const Grid &g, int b,
                     const T *pr, float *img, float w)
{
    const char *sl = static_cast<const char *>(store) + d.base + (size_t)j * d.slot;
    return launch_dct_decode(reinterpret_cast<const unsigned *>(sl + d.entries),
                             reinterpret_cast<const unsigned *>(sl + d.table), d.L,
                             static_cast<const BrickBasis *>(store), g, b, g.nz - b, pr, img, w,
                             (float *)nullptr, g_state.stream);
}

// The store's basis: the DCT-II (basis NULL) or the caller's (host: z[32][32],
// x[8][8], y[8][8], rows orthonormal to 1e-4, checked here).
int put_basis(void *store, const float *basis)
{
    if (!basis) {
        HIPCHK(dct_default_basis(static_cast<BrickBasis *>(store), g_state.stream), "basis");
        return RTMHIP_OK;
    }
    const float *m[3] = {basis, basis + 32 * 32, basis + 32 * 32 + 64};
    const int n[3] = {32, 8, 8};
    for (int a = 0; a < 3; ++a)
        for (int u = 0; u < n[a]; ++u)
            for (int v = 0; v < n[a]; ++v) {
                double dot = 0.0;
                for (int i = 0; i < n[a]; ++i)
                    dot += (double)m[a][u * n[a] + i] * (double)m[a][v * n[a] + i];
                if (std::fabs(dot - (u == v ? 1.0 : 0.0)) > 1e-4)
                    return fail(RTMHIP_ERR_INVALID, "basis rows are not orthonormal");
            }
    BrickBasis hb{};
    std::memcpy(&hb.z[0][0], basis, sizeof(hb.z));
    std::memcpy(&hb.x[0][0], basis + 32 * 32, sizeof(hb.x));
    std::memcpy(&hb.y[0][0], basis + 32 * 32 + 64, sizeof(hb.y));
    hb.learned = 1;
    HIPCHK(hipMemcpyAsync(store, &hb, sizeof(BrickBasis), hipMemcpyHostToDevice, g_state.stream),
           "hipMemcpyAsync");
    HIPCHK(hipStreamSynchronize(g_state.stream), "hipStreamSynchronize");
    return RTMHIP_OK;
}

int check_snap(const rtmhip_shot *shot, const void *snaps, int snap_format, int k)
{
    if (!snaps) return fail(RTMHIP_ERR_INVALID, "NULL snapshot buffer");
    if (snap_format != RTMHIP_SNAP_F32 && snap_format != RTMHIP_SNAP_F16)
        return fail(RTMHIP_ERR_UNSUPPORTED, "snapshot format %d (F32 or F16)", snap_format);
    if (k < 1 || k > shot->nt)
        return fail(RTMHIP_ERR_INVALID, "snapshot interval %d outside [1, nt]", k);
    if ((long long)shot->nxe * shot->nye >= (1LL << 31))
        return fail(RTMHIP_ERR_INVALID, "nxe*nye must stay below 2^31");
    return RTMHIP_OK;
}

}  // namespace

extern "C" {

int rtmhip_model_steps(const rtmhip_shot *shot, const float *V2DT2, const float *coefs,
                       const float *taper, const float *ricker, float **P, float **PP,
                       float *rcv_chunk, int it0, int nsteps)
{
    NEED_INIT();
    ShotCtx c;
    int rc;
    if ((rc = prepare_shot(shot, coefs, &c))) return rc;
    if ((rc = check_format(shot))) return rc;
    if (!V2DT2 || !taper || !ricker || !P || !PP || !*P || !*PP || !rcv_chunk)
        return fail(RTMHIP_ERR_INVALID, "NULL pointer");
    if (nsteps <= 0) return RTMHIP_OK;
    if ((rc = check_chunk(shot, it0, it0 + nsteps - 1))) return rc;
    return with_format(shot, [&](auto t) {
        using T = decltype(t);
        return model_loop<T>(shot, c, V2DT2, taper, ricker, as<T>(P), as<T>(PP), rcv_chunk, it0,
                             nsteps);
    });
}

int rtmhip_forward_steps(const rtmhip_shot *shot, int hbc, const float *V2DT2,
                         const float *coefs, const float *taper, const float *ricker, float **P,
                         float **PP, float *upb, int it0, int nsteps)
{
    NEED_INIT();
    ShotCtx c;
    int rc;
    if ((rc = prepare_shot(shot, coefs, &c))) return rc;
    if ((rc = check_format(shot))) return rc;
    if (!V2DT2 || !ricker || !P || !PP || !*P || !*PP || (hbc && (!taper || !upb)))
        return fail(RTMHIP_ERR_INVALID, "NULL pointer");
    if (hbc < 0 || hbc > 2) return fail(RTMHIP_ERR_INVALID, "hbc must be 0, 1 or 2");
    if (nsteps <= 0) return RTMHIP_OK;
    if (it0 < 0 || it0 + nsteps > shot->nt)
        return fail(RTMHIP_ERR_INVALID, "time range outside [0,%d)", shot->nt);
    // a one-chunk history holds the layers of one receiver chunk only
    if (hbc == 2 && (rc = check_chunk(shot, it0, it0 + nsteps - 1))) return rc;
    return with_format(shot, [&](auto t) {
        using T = decltype(t);
        return forward_loop<T>(shot, c, hbc, V2DT2, taper, ricker, as<T>(P), as<T>(PP),
                               reinterpret_cast<T *>(upb), it0, nsteps);
    });
}

int rtmhip_snap_buffer_bytes(const rtmhip_shot *shot, int snap_format, int k, size_t *bytes)
{
    if (!shot || !bytes) return fail(RTMHIP_ERR_INVALID, "NULL pointer");
    if (shot->nt <= 0 || shot->nze <= 2 * shot->blen)
        return fail(RTMHIP_ERR_INVALID, "bad shot geometry");
    if (k < 1) return fail(RTMHIP_ERR_INVALID, "snapshot interval %d below 1", k);
    const size_t es = snap_format == RTMHIP_SNAP_F32 ? 4 : snap_format == RTMHIP_SNAP_F16 ? 2 : 0;
    if (!es) return fail(RTMHIP_ERR_UNSUPPORTED, "snapshot format %d (F32 or F16)", snap_format);
    *bytes = (size_t)((shot->nt - 1) / k + 1) * snap_cells(shot) * es;
    return RTMHIP_OK;
}

int rtmhip_snap_forward_steps(const rtmhip_shot *shot, const float *V2DT2, const float *coefs,
                              const float *taper, const float *ricker, float **P, float **PP,
                              void *snaps, int snap_format, int k, int it0, int nsteps)
{
    NEED_INIT();
    ShotCtx c;
    int rc;
    if ((rc = prepare_shot(shot, coefs, &c))) return rc;
    if ((rc = check_format(shot))) return rc;
    if (!V2DT2 || !ricker || !taper || !P || !PP || !*P || !*PP)
        return fail(RTMHIP_ERR_INVALID, "NULL pointer");
    if ((rc = check_snap(shot, snaps, snap_format, k))) return rc;
    if (nsteps <= 0) return RTMHIP_OK;
    if (it0 < 0 || it0 + nsteps > shot->nt)
        return fail(RTMHIP_ERR_INVALID, "time range outside [0,%d)", shot->nt);
    return with_format(shot, [&](auto t) {
        using T = decltype(t);
        if (snap_format == RTMHIP_SNAP_F32)
            return snap_forward_loop<T, float>(shot, c, V2DT2, taper, ricker, as<T>(P), as<T>(PP),
                                               (float *)snaps, k, it0, nsteps);
        return snap_forward_loop<T, _Float16>(shot, c, V2DT2, taper, ricker, as<T>(P), as<T>(PP),
                                              (_Float16 *)snaps, k, it0, nsteps);
    });
}

int rtmhip_snap_backward_steps(const rtmhip_shot *shot, const float *V2DT2, const float *coefs,
                               const float *taper, float **Pr, float **PPr, const void *snaps,
                               int snap_format, int k, const float *rcv_chunk, float *img,
                               int it_hi, int nsteps)
{
    NEED_INIT();
    ShotCtx c;
    int rc;
    if ((rc = prepare_shot(shot, coefs, &c))) return rc;
    if ((rc = check_format(shot))) return rc;
    if (!V2DT2 || !taper || !Pr || !PPr || !*Pr || !*PPr || !rcv_chunk || !img)
        return fail(RTMHIP_ERR_INVALID, "NULL pointer");
    if ((rc = check_snap(shot, snaps, snap_format, k))) return rc;
    if (nsteps <= 0) return RTMHIP_OK;
    if ((rc = check_chunk(shot, it_hi - nsteps + 1, it_hi))) return rc;
    return with_format(shot, [&](auto t) {
        using T = decltype(t);
        if (snap_format == RTMHIP_SNAP_F32)
            return snap_backward_loop<T, float>(shot, c, V2DT2, taper, as<T>(Pr), as<T>(PPr),
                                                (const float *)snaps, k, rcv_chunk, img, it_hi,
                                                nsteps);
        return snap_backward_loop<T, _Float16>(shot, c, V2DT2, taper, as<T>(Pr), as<T>(PPr),
                                               (const _Float16 *)snaps, k, rcv_chunk, img, it_hi,
                                               nsteps);
    });
}

int rtmhip_snap_dct_sizes(const rtmhip_shot *shot, int k, int keep_div, size_t *store_bytes,
                          size_t *work_bytes)
{
    if (!shot) return fail(RTMHIP_ERR_INVALID, "NULL pointer");
    int rc = check_grid(shot->nxe, shot->nye, shot->nze);
    if (rc) return rc;
    if (shot->nt <= 0 || shot->nze <= 2 * shot->blen)
        return fail(RTMHIP_ERR_INVALID, "bad shot geometry");
    if (k < 1) return fail(RTMHIP_ERR_INVALID, "snapshot interval %d below 1", k);
    DctSlots d;
    if ((rc = dct_slots(dense_grid(shot->nxe, shot->nye, shot->nze), shot->blen, keep_div, &d)))
        return rc;
    if (store_bytes) *store_bytes = d.base + (size_t)((shot->nt - 1) / k + 1) * d.slot;
    if (work_bytes) *work_bytes = d.work;
    return RTMHIP_OK;
}

int rtmhip_snap_dct_forward_steps(const rtmhip_shot *shot, const float *V2DT2,
                                  const float *coefs, const float *taper, const float *ricker,
                                  float **P, float **PP, void *store, void *work,
                                  const float *basis, int k, int keep_div, int it0, int nsteps)
{
    NEED_INIT();
    ShotCtx c;
    int rc;
    if ((rc = prepare_shot(shot, coefs, &c))) return rc;
    if ((rc = check_format(shot))) return rc;
    if (!V2DT2 || !ricker || !taper || !P || !PP || !*P || !*PP || !store || !work)
        return fail(RTMHIP_ERR_INVALID, "NULL pointer");
    if (k < 1 || k > shot->nt)
        return fail(RTMHIP_ERR_INVALID, "snapshot interval %d outside [1, nt]", k);
    DctSlots d;
    if ((rc = dct_slots(c.g, shot->blen, keep_div, &d))) return rc;
    if (nsteps <= 0) return RTMHIP_OK;
    if (it0 < 0 || it0 + nsteps > shot->nt)
        return fail(RTMHIP_ERR_INVALID, "time range outside [0,%d)", shot->nt);
    if ((rc = put_basis(store, basis))) return rc;
    return with_format(shot, [&](auto t) {
        using T = decltype(t);
        return snap_forward_run(shot, c, V2DT2, taper, ricker, as<T>(P), as<T>(PP), k, it0,
                                nsteps, [&](const T *f, int j) {
                                    HIPCHK(dct_store(f, c.g, shot->blen, d, store, j, work),
                                           "snapshot transform coding");
                                    return RTMHIP_OK;
                                });
    });
}

int rtmhip_snap_dct_backward_steps(const rtmhip_shot *shot, const float *V2DT2,
                                   const float *coefs, const float *taper, float **Pr,
                                   float **PPr, const void *store, int k, int keep_div,
                                   const float *rcv_chunk, float *img, int it_hi, int nsteps)
{
    NEED_INIT();
    ShotCtx c;
    int rc;
    if ((rc = prepare_shot(shot, coefs, &c))) return rc;
    if ((rc = check_format(shot))) return rc;
    if (!V2DT2 || !taper || !Pr || !PPr || !*Pr || !*PPr || !rcv_chunk || !img || !store)
        return fail(RTMHIP_ERR_INVALID, "NULL pointer");
    if (k < 1 || k > shot->nt)
        return fail(RTMHIP_ERR_INVALID, "snapshot interval %d outside [1, nt]", k);
    DctSlots d;
    if ((rc = dct_slots(c.g, shot->blen, keep_div, &d))) return rc;
    if (nsteps <= 0) return RTMHIP_OK;
    if ((rc = check_chunk(shot, it_hi - nsteps + 1, it_hi))) return rc;
    return with_format(shot, [&](auto t) {
        using T = decltype(t);
        return snap_backward_run(shot, c, V2DT2, taper, as<T>(Pr), as<T>(PPr), k, rcv_chunk,
                                 it_hi, nsteps, [&](const T *pr, int j) {
                                     HIPCHK(dct_image(store, j, d, c.g, shot->blen, pr, img,
                                                      (float)k),
                                            "snapshot decoding");
                                     return RTMHIP_OK;
                                 });
    });
}

int rtmhip_dct_roundtrip(const void *field, int wave_format, int nxe, int nye, int nze,
                         int blen, int keep_div, void *store, void *work, const float *basis,
                         float *out, unsigned *kept)
{
    NEED_INIT();
    if (!field || !store || !work || !out) return fail(RTMHIP_ERR_INVALID, "NULL pointer");
    int rc = check_grid(nxe, nye, nze);
    if (rc) return rc;
    if (blen < 0 || 2 * blen >= nze) return fail(RTMHIP_ERR_INVALID, "bad blen %d", blen);
    const Grid g = dense_grid(nxe, nye, nze);
    DctSlots d;
    if ((rc = dct_slots(g, blen, keep_div, &d))) return rc;
    if ((rc = put_basis(store, basis))) return rc;
    const char *sl = static_cast<const char *>(store) + d.base;
    hipError_t e;
    switch (wave_format) {
    case RTMHIP_WAVE_F32: e = dct_store((const float *)field, g, blen, d, store, 0, work); break;
    case RTMHIP_WAVE_F16: e = dct_store((const _Float16 *)field, g, blen, d, store, 0, work); break;
    case RTMHIP_WAVE_BF16: e = dct_store((const __bf16 *)field, g, blen, d, store, 0, work); break;
    default: return fail(RTMHIP_ERR_UNSUPPORTED, "wave_format %d", wave_format);
    }
    HIPCHK(e, "transform coding");
    HIPCHK(launch_dct_decode(reinterpret_cast<const unsigned *>(sl + d.entries),
                             reinterpret_cast<const unsigned *>(sl + d.table), d.L,
                             static_cast<const BrickBasis *>(store), g, blen, nze - blen,
                             (const float *)nullptr, (float *)nullptr, 0.0f, out,
                             g_state.stream),
           "decoding");
    if (kept) {
        unsigned meta[3];
        HIPCHK(hipMemcpyAsync(meta, sl + d.meta, sizeof meta, hipMemcpyDeviceToHost,
                              g_state.stream),
               "hipMemcpyAsync");
        HIPCHK(hipStreamSynchronize(g_state.stream), "hipStreamSynchronize");
        *kept = meta[1];
    }
    return RTMHIP_OK;
}

int rtmhip_backward_steps(const rtmhip_shot *shot, int hbc, const float *V2DT2,
                          const float *coefs, const float *taper, float **Ps, float **PPs,
                          float **Pr, float **PPr, const float *upb, const float *rcv_chunk,
                          float *img, int it_hi, int nsteps, int fused)
{
    NEED_INIT();
    ShotCtx c;
    int rc;
    if ((rc = prepare_shot(shot, coefs, &c))) return rc;
    if ((rc = check_format(shot))) return rc;
    if (!V2DT2 || !Ps || !PPs || !Pr || !PPr || !*Ps || !*PPs || !*Pr || !*PPr || !rcv_chunk ||
        !img || (hbc && (!taper || !upb)))
        return fail(RTMHIP_ERR_INVALID, "NULL pointer");
    if (hbc < 0 || hbc > 2) return fail(RTMHIP_ERR_INVALID, "hbc must be 0, 1 or 2");
    if (nsteps <= 0) return RTMHIP_OK;
    if ((rc = check_chunk(shot, it_hi - nsteps + 1, it_hi))) return rc;
    return with_format(shot, [&](auto t) {
        using T = decltype(t);
        return backward_loop<T>(shot, c, hbc, V2DT2, taper, as<T>(Ps), as<T>(PPs), as<T>(Pr),
                                as<T>(PPr), reinterpret_cast<const T *>(upb), rcv_chunk, img,
                                it_hi, nsteps, fused);
    });
}

int rtmhip_set_tiling(int kernel_class, int tz, int by, int ry, int xchunk)
{
    NEED_INIT();
    if (kernel_class < 0 || kernel_class > 3)
        return fail(RTMHIP_ERR_INVALID, "kernel class %d not in 0..3", kernel_class);
    if (tz == 0) {
        g_state.tiling[kernel_class] = StencilTiling{};
        ++g_state.tiling_gen;
        return RTMHIP_OK;
    }
    const bool ok = kernel_class == 3 ? (tz == 64 && by == 1 && ry == 4)
                                      : stencil_tiling_supported(tz, by, ry);
    if (!ok || xchunk < 0)
        return fail(RTMHIP_ERR_UNSUPPORTED, "tile shape %dx%dx%d not compiled", tz, by, ry);
    g_state.tiling[kernel_class] = StencilTiling{tz, by, ry, xchunk};
    ++g_state.tiling_gen;
    return RTMHIP_OK;
}

int rtmhip_set_taper_fusion(int wave_format, int mode)
{
    NEED_INIT();
    if (wave_format < RTMHIP_WAVE_F32 || wave_format > RTMHIP_WAVE_BF16)
        return fail(RTMHIP_ERR_INVALID, "unknown wave_format %d", wave_format);
    if (mode < -1 || mode > 4)
        return fail(RTMHIP_ERR_INVALID,
                    "taper fusion mode %d (4 forward and modeling tapers on the output, 3 "
                    "modeling tapers on the output, 2 modeling tapers fused on load, 1 upper "
                    "tapers fused, 0 separate, -1 default)", mode);
    g_state.taper_mode[wave_format] = mode;
    g_state.taper_sep[wave_format] = mode < 0 ? kTaperSepDefault[wave_format] : mode == 0;
    g_state.taper_all_mode[wave_format] =
        (mode < 0 || mode == 3 || mode == 4) ? 2 : (mode == 2 ? 1 : 0);
    g_state.taper_outu[wave_format] = mode < 0 ? kTaperOutDefault[wave_format] : mode == 4;
    return RTMHIP_OK;
}

int rtmhip_get_taper_fusion(int wave_format, int *fused)
{
    NEED_INIT();
    if (wave_format < RTMHIP_WAVE_F32 || wave_format > RTMHIP_WAVE_BF16 || !fused)
        return fail(RTMHIP_ERR_INVALID, "bad arguments");
    *fused = g_state.taper_mode[wave_format];
    return RTMHIP_OK;
}

int rtmhip_get_tiling(int kernel_class, int *tz, int *by, int *ry, int *xchunk)
{
    NEED_INIT();
    if (kernel_class < 0 || kernel_class > 3 || !tz || !by || !ry || !xchunk)
        return fail(RTMHIP_ERR_INVALID, "bad arguments");
    const StencilTiling &t = g_state.tiling[kernel_class];
    *tz = t.tz;
    *by = t.by;
    *ry = t.ry;
    *xchunk = t.xchunk;
    return RTMHIP_OK;
}

int rtmhip_timing_enable(int on)
{
    NEED_INIT();
    g_state.timing = on != 0;
    return RTMHIP_OK;
}

int rtmhip_timing_read(int kernel_class, long long *launches, double *total_ms)
{
    NEED_INIT();
    if (kernel_class < 0 || kernel_class > 3 || !launches || !total_ms)
        return fail(RTMHIP_ERR_INVALID, "bad arguments");
    HIPCHK(hipStreamSynchronize(g_state.stream), "hipStreamSynchronize");
    Timing &t = g_state.tm[kernel_class];
    double sum = 0.0;
    for (size_t i = 0; i < t.used; ++i) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, t.start[i], t.stop[i]), "hipEventElapsedTime");
        sum += ms;
    }
    *launches = (long long)t.used;
    *total_ms = sum;
    t.used = 0;
    return RTMHIP_OK;
}

}  // extern "C"

// ===========================================================================
// Batches of shots.  Every shot keeps its own grid and buffers; a step of the
// batch is one launch per stencil class covering every shot (or one launch
// per shot, per class, with the class tiling) plus one launch of the small
// per-step operators of every shot.  Runs of steps replay a captured HIP
// graph whose small operators read the time index from device memory, so
// one captured sequence serves every time of the run.
// ===========================================================================

struct rtmhip_batch {
    int n = 0;
    int wave_format = 0;
    int blen = 0, nt = 0, hf = 0, radius = 0;
    Coefs cf{};
    const float *taper = nullptr;
    std::vector<rtmhip_shot> shots_v;
    std::vector<BatchShot> hs_v[2];  // host tables of a created batch: wave rows 6 and 8
    const rtmhip_shot *shots = nullptr;        // views of the tables above, or of
    BatchShot *hs[2] = {nullptr, nullptr};     // fixed storage (no-allocation runs)
    BatchShot *ds[2] = {nullptr, nullptr};
    long long nblk[2] = {0, 0};
    int fwd_blocks = 0, bwd_blocks = 0;
    int *ctr = nullptr;             // device time index of graph replays
    int cs = 0, cr = 0;             // P = f[c], PP = f[c ^ 1]
    int mode[3] = {0, 0, 0};        // per stencil class: 0 per-shot launches, 6/8 batched
    int graph_steps = 8;            // steps per captured graph (0: no graphs)
    bool fout = false;              // forward steps: upper taper on the output (this call)
    std::map<int, hipGraphExec_t> graphs;
    int tiling_gen = 0;             // g_state.tiling_gen when the graphs were captured
};

namespace {

int batch_tab(const rtmhip_batch *b, int cls) { return b->mode[cls] == 6 ? 0 : 1; }

// The forward steps of a batch store the upper taper of their output (as
// forward_loop) when the format asks for it, the forward stencil runs as
// per-shot launches, and every source lies at z >= blen.
bool batch_fwd_out(const rtmhip_batch *b)
{
    if (!g_state.taper_outu[b->wave_format] || b->mode[2] || b->radius != 4) return false;
    for (int i = 0; i < b->n; ++i) {
        const BatchShot &h = b->hs[0][i];
        if (h.src_off % h.g.sy < b->blen) return false;
    }
    return true;
}

// One forward step (HBC) of every shot at time it: with itp != nullptr the
// small operators read it from *itp (plus itoff).  form 0: tapers on load
// (taps as in forward_loop); 1: the upper taper on the output (first: P
// tapered in place before the step); 2: the last step of a call with form 1
// (PP tapered in place, plain step; first: P as well).
template <typename T>
int batch_fwd_step(rtmhip_batch *b, int taps, const int *itp, int itoff, int form = 0,
                   bool first = false)
{
    hipStream_t s = g_state.stream;
    b->cs ^= 1;
    const TaperFuse tf{b->taper, b->blen, taps};
    const int cls = 2;
    if (form) {
        for (int i = 0; i < b->n; ++i) {
            const BatchShot &h = b->hs[0][i];
            T *P = (T *)h.fs[b->cs], *PP = (T *)h.fs[b->cs ^ 1];
            if (first && form == 2)
                HIPCHK(launch_taper_upper2<T>(P, PP, b->taper, h.g, b->blen, s), "taper");
            else if (first)
                HIPCHK(launch_taper<T>(P, b->taper, h.g, b->blen, true, s), "taper");
            else if (form == 2)
                HIPCHK(launch_taper<T>(PP, b->taper, h.g, b->blen, true, s), "taper");
        }
        const TaperFuse to{b->taper, b->blen, 1, 3, b->hf};
        for (int i = 0; i < b->n; ++i) {
            const BatchShot &h = b->hs[0][i];
            HIPCHK(launch_stencil<T>(b->radius, (const T *)h.fs[b->cs], (T *)h.fs[b->cs ^ 1], h.V,
                                     b->cf, h.g, nullptr, form == 1 ? &to : nullptr,
                                     g_state.flag, s, class_tiling(form == 1 ? cls : 0)),
                   "stencil");
        }
        HIPCHK(launch_tail_fwd<T>(b->ds[0], b->n, b->fwd_blocks, b->cs, b->taper, b->blen,
                                  b->radius, b->hf, form == 1 ? 2 : 0, itp, itoff, s,
                                  form == 1),
               "forward tail");
        return RTMHIP_OK;
    }
    if (b->mode[cls]) {
        const int k = batch_tab(b, cls);
        HIPCHK(launch_stencil_batch<T>(b->mode[cls], b->ds[k], b->hs[k], b->n, b->nblk[k],
                                       false, b->cs, 0, b->cf, &tf, g_state.flag, s),
               "batched stencil");
    } else {
        for (int i = 0; i < b->n; ++i) {
            const BatchShot &h = b->hs[0][i];
            HIPCHK(launch_stencil<T>(b->radius, (const T *)h.fs[b->cs], (T *)h.fs[b->cs ^ 1], h.V,
                                     b->cf, h.g, nullptr, &tf, g_state.flag, s,
                                     class_tiling(cls)),
                   "stencil");
        }
    }
    HIPCHK(launch_tail_fwd<T>(b->ds[0], b->n, b->fwd_blocks, b->cs, b->taper, b->blen, b->radius,
                              b->hf, taps, itp, itoff, s),
           "forward tail");
    return RTMHIP_OK;
}

// One backward step (HBC, it < nt-2) of every shot: source reconstruction,
// receiver step fused with imaging, then the small operators.
template <typename T>
int batch_bwd_step(rtmhip_batch *b, int taps, const int *itp, int itoff)
{
    hipStream_t s = g_state.stream;
    const TaperFuse tf{b->taper, b->blen, taps};
    if (b->mode[0]) {
        const int k = batch_tab(b, 0);
        HIPCHK(launch_stencil_batch<T>(b->mode[0], b->ds[k], b->hs[k], b->n, b->nblk[k],
                                       false, b->cs, 0, b->cf, nullptr, g_state.flag, s),
               "batched stencil");
    } else {
        for (int i = 0; i < b->n; ++i) {
            const BatchShot &h = b->hs[0][i];
            HIPCHK(launch_stencil<T>(b->radius, (const T *)h.fs[b->cs], (T *)h.fs[b->cs ^ 1], h.V,
                                     b->cf, h.g, nullptr, nullptr, g_state.flag, s,
                                     class_tiling(0)),
                   "stencil");
        }
    }
    if (b->mode[1]) {
        const int k = batch_tab(b, 1);
        HIPCHK(launch_stencil_batch<T>(b->mode[1], b->ds[k], b->hs[k], b->n, b->nblk[k],
                                       true, b->cr, b->cs ^ 1, b->cf, &tf, g_state.flag, s),
               "batched stencil");
    } else {
        for (int i = 0; i < b->n; ++i) {
            const BatchShot &h = b->hs[0][i];
            ImageFuse fz{};
            fz.img = h.img;
            fz.ps = h.fs[b->cs ^ 1];
            fz.z0 = b->blen;
            fz.z1 = h.g.nz - b->blen;
            fz.rz = h.rz;
            fz.rx0 = h.rx0;
            fz.rdx = h.rdx;
            fz.rnx = h.rnx;
            fz.ry0 = h.ry0;
            fz.rdy = h.rdy;
            fz.rny = h.rny;
            HIPCHK(launch_stencil<T>(b->radius, (const T *)h.fr[b->cr], (T *)h.fr[b->cr ^ 1], h.V,
                                     b->cf, h.g, &fz, &tf, g_state.flag, s, class_tiling(1)),
                   "stencil");
        }
    }
    HIPCHK(launch_tail_bwd<T>(b->ds[0], b->n, b->bwd_blocks, b->cs, b->cr, b->taper, b->blen,
                              b->radius, b->hf, taps, b->nt, itp, itoff, s),
           "backward tail");
    b->cs ^= 1;
    b->cr ^= 1;
    return RTMHIP_OK;
}

// Captured sequence of graph_steps steps (forward dir = +1, backward -1)
// starting from the current parities; they return to those parities, so one
// executable graph per (direction, parities) serves every replay.
template <typename T>
int batch_graph(rtmhip_batch *b, int dir, hipGraphExec_t *out)
{
    if (b->tiling_gen != g_state.tiling_gen) {
        // per-shot launches in the captured graphs use the class tilings of
        // their capture: recapture after rtmhip_set_tiling
        HIPCHK(hipStreamSynchronize(g_state.stream), "hipStreamSynchronize");
        for (auto &kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
        b->graphs.clear();
        b->tiling_gen = g_state.tiling_gen;
    }
    const int key = (dir > 0 ? 0 : 4) | (b->cs << 1) | b->cr | (dir > 0 && b->fout ? 8 : 0);
    auto f = b->graphs.find(key);
    if (f != b->graphs.end()) {
        *out = f->second;
        return RTMHIP_OK;
    }
    hipStream_t s = g_state.stream;
    const int cs = b->cs, cr = b->cr;
    // Every launch shape of the captured steps was used by the eager first
    // step of this call (batch_run), so the launchers' occupancy cache
    // already holds it; the guard makes a miss an error instead of a runtime
    // query inside the capture.
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    capture_guard(true);
    int rc = RTMHIP_OK;
    for (int k = 0; k < b->graph_steps && rc == RTMHIP_OK; ++k)
        rc = dir > 0 ? batch_fwd_step<T>(b, 2, b->ctr, k, b->fout ? 1 : 0)
                     : batch_bwd_step<T>(b, 2, b->ctr, -k);
    hipError_t e = rc == RTMHIP_OK ? launch_tick(b->ctr, dir * b->graph_steps, false, s)
                                   : hipSuccess;
    capture_guard(false);
    hipGraph_t graph = nullptr;
    hipError_t e2 = hipStreamEndCapture(s, &graph);
    b->cs = cs;
    b->cr = cr;
    if (rc != RTMHIP_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return hip_fail(e, "graph capture");
    }
    HIPCHK(e2, "hipStreamEndCapture");
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIPCHK(e, "hipGraphInstantiate");
    b->graphs[key] = exec;
    *out = exec;
    return RTMHIP_OK;
}

// Steps it0 .. it0+nsteps-1 (dir +1) or it0 down to it0-nsteps+1 (dir -1)
// of the batched loop; the first step of a call tapers its PP once, later
// steps twice (the pending in-place taper of the previous step).
template <typename T>
int batch_run(rtmhip_batch *b, int dir, int it0, int nsteps)
{
    hipStream_t s = g_state.stream;
    // forward steps with the upper taper on the output: the first tapers P
    // in place, the last (never inside a graph) runs the reference's order
    b->fout = dir > 0 && batch_fwd_out(b);
    auto step = [&](int taps, int it, int k) {
        if (dir > 0 && b->fout)
            return batch_fwd_step<T>(b, taps, nullptr, it, k == nsteps - 1 ? 2 : 1, k == 0);
        return dir > 0 ? batch_fwd_step<T>(b, taps, nullptr, it)
                       : batch_bwd_step<T>(b, taps, nullptr, it);
    };
    int rc = step(1, it0, 0);
    if (rc) return rc;
    int done = 1;
    const int G = b->graph_steps;
    const int keep = b->fout ? 1 : 0;
    if (G > 0 && nsteps - keep - done >= G) {
        hipGraphExec_t exec;
        if ((rc = batch_graph<T>(b, dir, &exec))) return rc;
        HIPCHK(launch_tick(b->ctr, it0 + dir * done, true, s), "time index");
        while (nsteps - keep - done >= G) {
            HIPCHK(hipGraphLaunch(exec, s), "hipGraphLaunch");
            done += G;  // parities unchanged: G is even
        }
    }
    for (; done < nsteps; ++done)
        if ((rc = step(2, it0 + dir * done, done))) return rc;
    if (dir > 0 && b->fout) return RTMHIP_OK;
    // the pending in-place taper: P of the source field (forward) or the
    // older receiver field (backward), as the per-shot loops leave them
    for (int i = 0; i < b->n; ++i) {
        const BatchShot &h = b->hs[0][i];
        T *f = dir > 0 ? (T *)h.fs[b->cs] : (T *)h.fr[b->cr ^ 1];
        HIPCHK(launch_taper<T>(f, b->taper, h.g, b->blen, true, s), "taper");
    }
    return RTMHIP_OK;
}

template <typename T>
int batch_backward(rtmhip_batch *b, int it_hi, int nsteps)
{
    // the two seeding steps (it = nt-1, nt-2: the source field is the
    // forward's last two states, not reconstructed) run per shot
    int it = it_hi;
    int left = nsteps;
    const int seed = std::min(left, std::max(0, it - (b->nt - 3)));
    if (seed > 0) {
        for (int i = 0; i < b->n; ++i) {
            BatchShot &h = b->hs[0][i];
            ShotCtx c;
            c.g = h.g;
            c.radius = b->radius;
            c.hf = b->hf;
            c.cf = b->cf;
            c.src_off = h.src_off;
            c.rx0 = h.rx0;
            c.ry0 = h.ry0;
            c.rz = h.rz;
            T *Ps = (T *)h.fs[b->cs], *PPs = (T *)h.fs[b->cs ^ 1];
            T *Pr = (T *)h.fr[b->cr], *PPr = (T *)h.fr[b->cr ^ 1];
            int rc = backward_loop<T>(&b->shots[i], c, 1, h.V, b->taper, &Ps, &PPs, &Pr, &PPr,
                                      (const T *)h.upb, h.rcv, h.img, it, seed, 1);
            if (rc) return rc;
            if (i == b->n - 1) {
                b->cs = Ps == (T *)h.fs[0] ? 0 : 1;
                b->cr = Pr == (T *)h.fr[0] ? 0 : 1;
            }
        }
        it -= seed;
        left -= seed;
    }
    if (left <= 0) return RTMHIP_OK;
    return batch_run<T>(b, -1, it, left);
}

// Validates a batch's shots and fills the host table hs (n entries, the
// 6-row plan not yet made), the shot array sh and the batch's shared
// fields; allocates nothing.
int batch_fill(rtmhip_batch *b, const rtmhip_batch_shot *shots, int nshots, const float *coefs,
               const float *taper, BatchShot *hs, rtmhip_shot *sh)
{
    b->n = nshots;
    const rtmhip_shot &s0 = shots[0].shot;
    for (int i = 0; i < nshots; ++i) {
        const rtmhip_batch_shot &bs = shots[i];
        const rtmhip_shot &shot = bs.shot;
        ShotCtx c;
        int rc;
        if ((rc = prepare_shot(&shot, coefs, &c))) return rc;
        if ((rc = check_format(&shot))) return rc;
        if (shot.order != 8) return fail(RTMHIP_ERR_UNSUPPORTED, "batches need stencil order 8");
        if (shot.blen != s0.blen || shot.nt != s0.nt || shot.wave_format != s0.wave_format)
            return fail(RTMHIP_ERR_INVALID, "shot %d: blen, nt and wave_format must match shot 0",
                        i);
        if (shot.blen > 8192)
            return fail(RTMHIP_ERR_UNSUPPORTED, "shot %d: blen %d above 8192", i, shot.blen);
        if (shot.ntstep != shot.nt)
            return fail(RTMHIP_ERR_INVALID, "shot %d: batches keep whole records (ntstep == nt)",
                        i);
        if (!bs.V2DT2 || !bs.ricker || !bs.fs[0] || !bs.fs[1] || !bs.fr[0] || !bs.fr[1] ||
            !bs.upb || !bs.rcv || !bs.img)
            return fail(RTMHIP_ERR_INVALID, "shot %d: NULL pointer", i);
        BatchShot h{};
        h.fs[0] = bs.fs[0];
        h.fs[1] = bs.fs[1];
        h.fr[0] = bs.fr[0];
        h.fr[1] = bs.fr[1];
        h.V = bs.V2DT2;
        h.img = bs.img;
        h.upb = bs.upb;
        h.rcv = bs.rcv;
        h.ricker = bs.ricker;
        h.g = c.g;
        h.src_off = c.src_off;
        h.rx0 = c.rx0;
        h.ry0 = c.ry0;
        h.rz = c.rz;
        h.rdx = std::max(1, shot.rcv.dist_x);
        h.rdy = std::max(1, shot.rcv.dist_y);
        h.rnx = shot.rcv.count_x;
        h.rny = shot.rcv.count_y;
        hs[i] = h;
        sh[i] = shot;
        if (i == 0) {
            b->radius = c.radius;
            b->hf = c.hf;
            b->cf = c.cf;
        }
    }
    b->wave_format = s0.wave_format;
    b->blen = s0.blen;
    b->nt = s0.nt;
    b->taper = taper;
    long long fw = 0, bw = 0;
    for (int i = 0; i < nshots; ++i) {
        BatchShot &h = hs[i];
        const long long nf = tail_fwd_items(h.g, b->blen, b->radius, b->hf);
        const long long nb = tail_bwd_items(h.g, b->blen, b->radius, b->hf, h.rnx, h.rny);
        if (nf >= (1LL << 31) || nb >= (1LL << 31) ||
            (long long)h.g.nx * h.g.ny * b->hf >= (1LL << 31))
            return fail(RTMHIP_ERR_INVALID, "a shot's border work exceeds 2^31 cells");
        h.fwd0 = (int)fw;
        h.bwd0 = (int)bw;
        h.nfwd = (int)nf;
        h.nbwd = (int)nb;
        fw += tail_blocks(nf);
        bw += tail_blocks(nb);
        if (fw >= (1LL << 31) || bw >= (1LL << 31))
            return fail(RTMHIP_ERR_INVALID, "batch too large");
    }
    b->fwd_blocks = (int)fw;
    b->bwd_blocks = (int)bw;
    return RTMHIP_OK;
}

}  // namespace

extern "C" {

int rtmhip_batch_create(rtmhip_batch **out, const rtmhip_batch_shot *shots, int nshots,
                        const float *coefs, const float *taper)
{
    NEED_INIT();
    if (!out || !shots || !coefs || !taper) return fail(RTMHIP_ERR_INVALID, "NULL pointer");
    *out = nullptr;
    if (nshots <= 0) return fail(RTMHIP_ERR_INVALID, "empty batch");
    auto b = std::make_unique<rtmhip_batch>();
    b->shots_v.resize(nshots);
    b->hs_v[0].resize(nshots);
    int rc = batch_fill(b.get(), shots, nshots, coefs, taper, b->hs_v[0].data(),
                        b->shots_v.data());
    if (rc) return rc;
    b->hs_v[1] = b->hs_v[0];
    b->shots = b->shots_v.data();
    b->hs[0] = b->hs_v[0].data();
    b->hs[1] = b->hs_v[1].data();
    b->tiling_gen = g_state.tiling_gen;
    b->nblk[0] = plan_stencil_batch(6, b->hs[0], nshots, b->radius, b->blen);
    b->nblk[1] = plan_stencil_batch(8, b->hs[1], nshots, b->radius, b->blen);
    if (b->nblk[0] <= 0 || b->nblk[1] <= 0)
        return fail(RTMHIP_ERR_INVALID, "a shot's grid leaves no interior");
    const size_t bytes = sizeof(BatchShot) * (size_t)nshots;
    for (int k = 0; k < 2; ++k) {
        void *p = nullptr;
        hipError_t e = hipMalloc(&p, bytes);
        if (e == hipSuccess) {
            b->ds[k] = (BatchShot *)p;
            e = hipMemcpyAsync(p, b->hs[k], bytes, hipMemcpyHostToDevice, g_state.stream);
        }
        if (e != hipSuccess) {
            rtmhip_batch_destroy(b.release());
            return hip_fail(e, "batch tables");
        }
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, sizeof(int));
    if (e != hipSuccess) {
        rtmhip_batch_destroy(b.release());
        return hip_fail(e, "hipMalloc");
    }
    b->ctr = (int *)p;
    e = hipStreamSynchronize(g_state.stream);  // the host tables are uploaded
    if (e != hipSuccess) {
        rtmhip_batch_destroy(b.release());
        return hip_fail(e, "hipStreamSynchronize");
    }
    *out = b.release();
    return RTMHIP_OK;
}

int rtmhip_batch_destroy(rtmhip_batch *b)
{
    if (!b) return RTMHIP_OK;
    if (g_state.init) (void)hipStreamSynchronize(g_state.stream);
    for (auto &kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
    for (int k = 0; k < 2; ++k)
        if (b->ds[k]) (void)hipFree(b->ds[k]);
    if (b->ctr) (void)hipFree(b->ctr);
    delete b;
    return RTMHIP_OK;
}

int rtmhip_batch_set_stencil(rtmhip_batch *b, int kernel_class, int mode)
{
    NEED_INIT();
    if (!b) return fail(RTMHIP_ERR_INVALID, "NULL batch");
    if (kernel_class < 0 || kernel_class > 2)
        return fail(RTMHIP_ERR_INVALID, "kernel class %d not in 0..2", kernel_class);
    if (mode != 0 && mode != 6 && mode != 8)
        return fail(RTMHIP_ERR_UNSUPPORTED, "batched stencil mode %d (0, 6 or 8)", mode);
    if (b->mode[kernel_class] != mode) {
        HIPCHK(hipStreamSynchronize(g_state.stream), "hipStreamSynchronize");
        for (auto &kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
        b->graphs.clear();
        b->mode[kernel_class] = mode;
    }
    return RTMHIP_OK;
}

int rtmhip_batch_set_graph(rtmhip_batch *b, int steps)
{
    NEED_INIT();
    if (!b) return fail(RTMHIP_ERR_INVALID, "NULL batch");
    if (steps < 0 || (steps & 1))
        return fail(RTMHIP_ERR_INVALID, "steps per graph must be 0 or even (got %d)", steps);
    if (b->graph_steps != steps) {
        HIPCHK(hipStreamSynchronize(g_state.stream), "hipStreamSynchronize");
        for (auto &kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
        b->graphs.clear();
        b->graph_steps = steps;
    }
    return RTMHIP_OK;
}

int rtmhip_batch_fields(const rtmhip_batch *b, int i, float **P, float **PP, float **Pr,
                        float **PPr)
{
    NEED_INIT();
    if (!b || i < 0 || i >= b->n) return fail(RTMHIP_ERR_INVALID, "bad batch or shot index");
    const BatchShot &h = b->hs[0][i];
    if (P) *P = (float *)h.fs[b->cs];
    if (PP) *PP = (float *)h.fs[b->cs ^ 1];
    if (Pr) *Pr = (float *)h.fr[b->cr];
    if (PPr) *PPr = (float *)h.fr[b->cr ^ 1];
    return RTMHIP_OK;
}

int rtmhip_batch_forward_steps(rtmhip_batch *b, int it0, int nsteps)
{
    NEED_INIT();
    if (!b) return fail(RTMHIP_ERR_INVALID, "NULL batch");
    if (nsteps <= 0) return RTMHIP_OK;
    if (it0 < 0 || it0 + nsteps > b->nt)
        return fail(RTMHIP_ERR_INVALID, "time range outside [0,%d)", b->nt);
    const rtmhip_shot *s = &b->shots[0];
    return with_format(s, [&](auto t) {
        using T = decltype(t);
        return batch_run<T>(b, +1, it0, nsteps);
    });
}

int rtmhip_batch_backward_steps(rtmhip_batch *b, int it_hi, int nsteps)
{
    NEED_INIT();
    if (!b) return fail(RTMHIP_ERR_INVALID, "NULL batch");
    if (nsteps <= 0) return RTMHIP_OK;
    if (it_hi >= b->nt || it_hi - nsteps + 1 < 0)
        return fail(RTMHIP_ERR_INVALID, "time range outside [0,%d)", b->nt);
    const rtmhip_shot *s = &b->shots[0];
    return with_format(s, [&](auto t) {
        using T = decltype(t);
        return batch_backward<T>(b, it_hi, nsteps);
    });
}

int rtmhip_batch_steps_noalloc(rtmhip_batch_shot *shots, int nshots, const float *coefs,
                               const float *taper, int direction, int it, int nsteps)
{
    NEED_INIT();
    if (!shots || !coefs || !taper) return fail(RTMHIP_ERR_INVALID, "NULL pointer");
    if (nshots <= 0) return fail(RTMHIP_ERR_INVALID, "empty batch");
    if (nshots > kNoAllocShots)
        return fail(RTMHIP_ERR_INVALID,
                    "rtmhip_batch_steps_noalloc: %d shots, the staging array holds at most %d",
                    nshots, kNoAllocShots);
    if (direction != 1 && direction != -1)
        return fail(RTMHIP_ERR_INVALID, "direction must be +1 (forward) or -1 (backward)");
    if (g_state.device < 0 || g_state.device >= kMaxDevices || !g_scratch[g_state.device])
        return fail(RTMHIP_ERR_NOT_INIT, "no staging array on device %d", g_state.device);
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    static BatchShot hs[kNoAllocShots];
    static rtmhip_shot sh[kNoAllocShots];
    rtmhip_batch b;     // its containers stay empty: nothing is allocated
    b.graph_steps = 0;  // eager steps (instantiating a graph allocates)
    int rc = batch_fill(&b, shots, nshots, coefs, taper, hs, sh);
    if (rc) return rc;
    if (nsteps <= 0) return RTMHIP_OK;
    if (direction > 0 ? (it < 0 || it + nsteps > b.nt) : (it >= b.nt || it - nsteps + 1 < 0))
        return fail(RTMHIP_ERR_INVALID, "time range outside [0,%d)", b.nt);
    BatchShot *d = g_scratch[g_state.device];
    b.shots = sh;
    b.hs[0] = b.hs[1] = hs;
    b.ds[0] = b.ds[1] = d;
    // stream-ordered after the previous call's kernels that read the array
    HIPCHK(hipMemcpyAsync(d, hs, sizeof(BatchShot) * (size_t)nshots, hipMemcpyHostToDevice,
                          g_state.stream),
           "staging the shot table");
    rc = with_format(&sh[0], [&](auto t) {
        using T = decltype(t);
        return direction > 0 ? batch_run<T>(&b, +1, it, nsteps) : batch_backward<T>(&b, it, nsteps);
    });
    if (rc) return rc;
    for (int i = 0; i < nshots; ++i) {
        if (b.cs) std::swap(shots[i].fs[0], shots[i].fs[1]);
        if (b.cr) std::swap(shots[i].fr[0], shots[i].fr[1]);
    }
    return RTMHIP_OK;
}

}  // extern "C"

// ===========================================================================
// Snapshot autoencoder: a 4-layer patch-convolution network that compresses
// 32^3 FP16 bricks of a source-wavefield snapshot 16x (2048 FP16 codes per
// brick plus one FP32 scale), trained on the device by its own forward and
// backward kernels (matrix-core GEMMs) with Adam.
// ===========================================================================

namespace {

struct AeLayerDef {
    int K, N, act;      // weights [K][N] row-major, then the bias [N]
    long long rows;     // GEMM rows per brick
};
constexpr AeLayerDef kAe[4] = {{8, 8, 1, 4096}, {512, 32, 0, 64}, {32, 512, 1, 64},
                               {8, 8, 0, 4096}};
constexpr long long kBrickCells = 32768, kCodes = 2048;

long long ae_param_offset(int layer)
{
    long long o = 0;
    for (int i = 0; i < layer; ++i) o += (long long)kAe[i].K * kAe[i].N + kAe[i].N;
    return o;
}
constexpr long long kAeParams = 8 * 8 + 8 + 512 * 32 + 32 + 32 * 512 + 512 + 8 * 8 + 8;

}  // namespace

struct rtmhip_ae {
    int maxb = 0;
    float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // kAeParams each
    void *x0 = nullptr;     // FP16 input bricks [maxb][32768]
    float *h1 = nullptr, *z = nullptr, *h3 = nullptr, *y = nullptr;
    float *d1 = nullptr, *dz = nullptr, *d3 = nullptr, *dy = nullptr;
    float *loss = nullptr;  // device accumulator
    long long steps = 0;    // Adam steps taken
};

namespace {

float *ae_w(const rtmhip_ae *a, int l) { return a->w + ae_param_offset(l); }
float *ae_b(const rtmhip_ae *a, int l) { return ae_w(a, l) + (long long)kAe[l].K * kAe[l].N; }
float *ae_gw(const rtmhip_ae *a, int l) { return a->g + ae_param_offset(l); }
float *ae_gb(const rtmhip_ae *a, int l) { return ae_gw(a, l) + (long long)kAe[l].K * kAe[l].N; }

// Row-major GEMM of one layer: out[rows][N] = act(in[rows][K] W + b).
hipError_t ae_layer(const rtmhip_ae *a, int l, const void *in, int in_half, float *out, int n)
{
    AeGemm p{};
    p.A = in;
    p.sam = kAe[l].K;
    p.sak = 1;
    p.a_half = in_half;
    p.B = ae_w(a, l);
    p.sbk = kAe[l].N;
    p.sbn = 1;
    p.C = out;
    p.ldc = kAe[l].N;
    p.bias = ae_b(a, l);
    p.M = kAe[l].rows * n;
    p.N = kAe[l].N;
    p.K = kAe[l].K;
    p.act = kAe[l].act;
    return launch_ae_gemm(p, g_state.stream);
}

hipError_t ae_forward(const rtmhip_ae *a, const void *x0, int n)
{
    hipError_t e;
    if ((e = ae_layer(a, 0, x0, 1, a->h1, n)) != hipSuccess) return e;
    if ((e = ae_layer(a, 1, a->h1, 0, a->z, n)) != hipSuccess) return e;
    if ((e = ae_layer(a, 2, a->z, 0, a->h3, n)) != hipSuccess) return e;
    return ae_layer(a, 3, a->h3, 0, a->y, n);
}

// Weight gradient of layer l: gW[K][N] += in^T (rows x K) * d (rows x N);
// input gradient din[rows][K] = d * W^T (skipped when din is null).
hipError_t ae_layer_back(const rtmhip_ae *a, int l, const void *in, int in_half, const float *d,
                         float *din, int n)
{
    hipStream_t s = g_state.stream;
    const long long rows = kAe[l].rows * n;
    const int K = kAe[l].K, N = kAe[l].N;
    AeGemm p{};
    p.A = in;
    p.sam = 1;
    p.sak = K;
    p.a_half = in_half;
    p.B = d;
    p.sbk = N;
    p.sbn = 1;
    p.C = ae_gw(a, l);
    p.ldc = N;
    p.M = K;
    p.N = N;
    p.K = rows;
    p.accumulate = 1;
    p.kchunk = 0;  // sized by the launcher
    hipError_t e = launch_ae_gemm(p, s);
    if (e != hipSuccess) return e;
    if ((e = launch_colsum(d, rows, N, ae_gb(a, l), s)) != hipSuccess) return e;
    if (!din) return hipSuccess;
    AeGemm q{};
    q.A = d;
    q.sam = N;
    q.sak = 1;
    q.B = ae_w(a, l);  // B(k, n) = W[n][k]
    q.sbk = 1;
    q.sbn = N;
    q.C = din;
    q.ldc = K;
    q.M = rows;
    q.N = K;
    q.K = N;
    return launch_ae_gemm(q, s);
}

// Gradient of 0.5 * sum (Y - X)^2 over n bricks into a->g (zeroed first);
// the sum of squares is added to *loss (device).
hipError_t ae_backward(const rtmhip_ae *a, const void *x0, int n, float *loss)
{
    hipStream_t s = g_state.stream;
    const long long cells = kBrickCells * n;
    hipError_t e;
    if ((e = launch_fill(a->g, (size_t)kAeParams, 0.f, s)) != hipSuccess) return e;
    if ((e = launch_ae_residual(a->y, x0, a->dy, cells, loss, s)) != hipSuccess) return e;
    if ((e = ae_layer_back(a, 3, a->h3, 0, a->dy, a->d3, n)) != hipSuccess) return e;
    if ((e = launch_tanh_back(a->d3, a->h3, cells, s)) != hipSuccess) return e;
    if ((e = ae_layer_back(a, 2, a->z, 0, a->d3, a->dz, n)) != hipSuccess) return e;
    if ((e = ae_layer_back(a, 1, a->h1, 0, a->dz, a->d1, n)) != hipSuccess) return e;
    if ((e = launch_tanh_back(a->d1, a->h1, cells, s)) != hipSuccess) return e;
    return ae_layer_back(a, 0, x0, 1, a->d1, nullptr, n);
}

uint64_t splitmix64(uint64_t &x)
{
    uint64_t z = (x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

int ae_check_region(int nxe, int nye, int nze, int x0, int y0, int z0, int x1, int y1, int z1,
                    AeRegion *r, Grid *g)
{
    int rc = check_grid(nxe, nye, nze);
    if (rc) return rc;
    if (x0 < 0 || y0 < 0 || z0 < 0 || x1 > nxe || y1 > nye || z1 > nze || x0 >= x1 || y0 >= y1 ||
        z0 >= z1)
        return fail(RTMHIP_ERR_INVALID, "region outside the grid or empty");
    r->x0 = x0;
    r->y0 = y0;
    r->z0 = z0;
    r->x1 = x1;
    r->y1 = y1;
    r->z1 = z1;
    r->nbx = (x1 - x0 + 31) / 32;
    r->nby = (y1 - y0 + 31) / 32;
    r->nbz = (z1 - z0 + 31) / 32;
    *g = dense_grid(nxe, nye, nze);
    return RTMHIP_OK;
}

// n FP16 bricks -> n x 2048 FP16 codes, in passes of at most maxb bricks.
hipError_t ae_encode(const rtmhip_ae *a, const void *bricks, int n, void *codes)
{
    hipError_t e = hipSuccess;
    for (int b0 = 0; b0 < n && e == hipSuccess; b0 += a->maxb) {
        const int nb = std::min(a->maxb, n - b0);
        const char *in = static_cast<const char *>(bricks) + (size_t)b0 * kBrickCells * 2;
        e = ae_layer(a, 0, in, 1, a->h1, nb);
        if (e == hipSuccess) e = ae_layer(a, 1, a->h1, 0, a->z, nb);
        if (e == hipSuccess)
            e = launch_f32_to_f16(a->z, static_cast<char *>(codes) + (size_t)b0 * kCodes * 2,
                                  kCodes * nb, g_state.stream);
    }
    return e;
}

// Codes + scales of the n bricks of region r -> the region's cells of field.
hipError_t ae_decode(const rtmhip_ae *a, const void *codes, const float *scales, int n,
                     const Grid &g, const AeRegion &r, float *field)
{
    hipError_t e = hipSuccess;
    for (int b0 = 0; b0 < n && e == hipSuccess; b0 += a->maxb) {
        const int nb = std::min(a->maxb, n - b0);
        e = launch_f16_to_f32(static_cast<const char *>(codes) + (size_t)b0 * kCodes * 2, a->z,
                              kCodes * nb, g_state.stream);
        if (e == hipSuccess) e = ae_layer(a, 2, a->z, 0, a->h3, nb);
        if (e == hipSuccess) e = ae_layer(a, 3, a->h3, 0, a->y, nb);
        if (e == hipSuccess) e = launch_unbricks(a->y, scales, g, r, field, b0, nb, g_state.stream);
    }
    return e;
}

// The imaging region of a shot's snapshots (all x, y; z in [z0, z0 + lz))
// as bricks: on the extended grid (z0 = blen) or on a compact
// [nxe][nye][nze-2*blen] field (z0 = 0).
AeRegion snap_region(const rtmhip_shot *shot, int z0)
{
    const int lz = shot->nze - 2 * shot->blen;
    AeRegion r;
    r.x0 = 0;
    r.y0 = 0;
    r.z0 = z0;
    r.x1 = shot->nxe;
    r.y1 = shot->nye;
    r.z1 = z0 + lz;
    r.nbx = (shot->nxe + 31) / 32;
    r.nby = (shot->nye + 31) / 32;
    r.nbz = (lz + 31) / 32;
    return r;
}

long long snap_region_bricks(const rtmhip_shot *shot)
{
    const AeRegion r = snap_region(shot, 0);
    return (long long)r.nbx * r.nby * r.nbz;
}

size_t snap_ae_work_bytes(const rtmhip_shot *shot)
{
    const size_t bricks = (size_t)snap_region_bricks(shot) * kBrickCells * 2;
    const size_t field = snap_cells(shot) * 4;
    return std::max(bricks, field);
}

int check_snap_ae(const rtmhip_shot *shot, const rtmhip_ae *a, const void *codes,
                  const float *scales, const void *work, int k)
{
    if (!a || !codes || !scales || !work) return fail(RTMHIP_ERR_INVALID, "NULL pointer");
    if (shot->wave_format != RTMHIP_WAVE_F32)
        return fail(RTMHIP_ERR_UNSUPPORTED, "coded snapshots need FP32 wavefields");
    if (snap_region_bricks(shot) > 0x7fffffffLL)
        return fail(RTMHIP_ERR_INVALID, "imaging region of more than 2^31 bricks");
    return check_snap(shot, codes, RTMHIP_SNAP_F32, k);
}

}  // namespace

extern "C" {

int rtmhip_ae_param_count(void) { return (int)kAeParams; }

int rtmhip_ae_create(rtmhip_ae **out, int max_bricks, unsigned long long seed)
{
    NEED_INIT();
    if (!out) return fail(RTMHIP_ERR_INVALID, "NULL pointer");
    *out = nullptr;
    if (max_bricks <= 0 || max_bricks > 4096)
        return fail(RTMHIP_ERR_INVALID, "max_bricks %d not in 1..4096", max_bricks);
    auto a = std::make_unique<rtmhip_ae>();
    a->maxb = max_bricks;
    const long long cells = kBrickCells * max_bricks, codes = kCodes * max_bricks;
    struct Alloc {
        void **p;
        size_t bytes;
    } allocs[] = {
        {(void **)&a->w, kAeParams * 4}, {(void **)&a->g, kAeParams * 4},
        {(void **)&a->m, kAeParams * 4}, {(void **)&a->v, kAeParams * 4},
        {&a->x0, (size_t)cells * 2},     {(void **)&a->h1, (size_t)cells * 4},
        {(void **)&a->z, (size_t)codes * 4}, {(void **)&a->h3, (size_t)cells * 4},
        {(void **)&a->y, (size_t)cells * 4}, {(void **)&a->d1, (size_t)cells * 4},
        {(void **)&a->dz, (size_t)codes * 4}, {(void **)&a->d3, (size_t)cells * 4},
        {(void **)&a->dy, (size_t)cells * 4}, {(void **)&a->loss, 4},
    };
    for (auto &al : allocs) {
        hipError_t e = hipMalloc(al.p, al.bytes);
        if (e != hipSuccess) {
            rtmhip_ae_destroy(a.release());
            return hip_fail(e, "hipMalloc");
        }
    }
    // Glorot-uniform weights from the seed, zero biases, zero Adam moments
    std::vector<float> w((size_t)kAeParams, 0.f);
    uint64_t st = seed;
    for (int l = 0; l < 4; ++l) {
        const float lim = std::sqrt(6.0f / (float)(kAe[l].K + kAe[l].N));
        float *p = w.data() + ae_param_offset(l);
        for (long long i = 0; i < (long long)kAe[l].K * kAe[l].N; ++i) {
            const double u = (double)(splitmix64(st) >> 11) * (1.0 / 9007199254740992.0);
            p[i] = (float)((2.0 * u - 1.0) * lim);
        }
    }
    hipError_t e = hipMemcpy(a->w, w.data(), kAeParams * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_fill(a->m, (size_t)kAeParams, 0.f, g_state.stream);
    if (e == hipSuccess) e = launch_fill(a->v, (size_t)kAeParams, 0.f, g_state.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g_state.stream);
    if (e != hipSuccess) {
        rtmhip_ae_destroy(a.release());
        return hip_fail(e, "autoencoder set-up");
    }
    *out = a.release();
    return RTMHIP_OK;
}

int rtmhip_ae_destroy(rtmhip_ae *a)
{
    if (!a) return RTMHIP_OK;
    if (g_state.init) (void)hipStreamSynchronize(g_state.stream);
    void *ps[] = {a->w, a->g, a->m, a->v, a->x0, a->h1, a->z, a->h3,
                  a->y, a->d1, a->dz, a->d3, a->dy, a->loss};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    delete a;
    return RTMHIP_OK;
}

int rtmhip_ae_get_params(const rtmhip_ae *a, float *host, int which)
{
    NEED_INIT();
    if (!a || !host || which < 0 || which > 1) return fail(RTMHIP_ERR_INVALID, "bad arguments");
    HIPCHK(hipMemcpyAsync(host, which ? a->g : a->w, kAeParams * 4, hipMemcpyDeviceToHost,
                          g_state.stream),
           "hipMemcpyAsync");
    HIPCHK(hipStreamSynchronize(g_state.stream), "hipStreamSynchronize");
    return RTMHIP_OK;
}

int rtmhip_ae_set_params(rtmhip_ae *a, const float *host)
{
    NEED_INIT();
    if (!a || !host) return fail(RTMHIP_ERR_INVALID, "bad arguments");
    HIPCHK(hipMemcpyAsync(a->w, host, kAeParams * 4, hipMemcpyHostToDevice, g_state.stream),
           "hipMemcpyAsync");
    HIPCHK(hipStreamSynchronize(g_state.stream), "hipStreamSynchronize");
    return RTMHIP_OK;
}

int rtmhip_ae_gradient(rtmhip_ae *a, const void *bricks, int n, float *loss)
{
    NEED_INIT();
    if (!a || !bricks || !loss) return fail(RTMHIP_ERR_INVALID, "NULL pointer");
    if (n <= 0 || n > a->maxb) return fail(RTMHIP_ERR_INVALID, "n %d not in 1..%d", n, a->maxb);
    HIPCHK(launch_fill(a->loss, 1, 0.f, g_state.stream), "fill");
    HIPCHK(ae_forward(a, bricks, n), "autoencoder forward");
    HIPCHK(ae_backward(a, bricks, n, a->loss), "autoencoder backward");
    HIPCHK(hipMemcpyAsync(loss, a->loss, 4, hipMemcpyDeviceToHost, g_state.stream), "copy");
    HIPCHK(hipStreamSynchronize(g_state.stream), "hipStreamSynchronize");
    *loss *= 0.5f;
    return RTMHIP_OK;
}

int rtmhip_ae_train(rtmhip_ae *a, const void *bricks, int nbricks, int iters, int batch, float lr,
                    unsigned long long seed, float *losses)
{
    NEED_INIT();
    if (!a || !bricks) return fail(RTMHIP_ERR_INVALID, "NULL pointer");
    if (nbricks <= 0 || iters <= 0 || batch <= 0 || batch > a->maxb || batch > nbricks)
        return fail(RTMHIP_ERR_INVALID, "bad sizes");
    if ((long long)iters * batch > (1LL << 28)) return fail(RTMHIP_ERR_INVALID, "too many iterations");
    hipStream_t s = g_state.stream;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    // every iteration's batch, drawn without replacement from reshuffled
    // passes over the bricks (Fisher-Yates on a seeded splitmix64 stream)
    uint64_t st = seed;
    std::vector<int> perm(nbricks), picks((size_t)iters * batch);
    for (int i = 0; i < nbricks; ++i) perm[i] = i;
    size_t pos = (size_t)nbricks;
    for (auto &p : picks) {
        if (pos >= (size_t)nbricks) {
            for (int i = nbricks - 1; i > 0; --i) {
                const int k = (int)(splitmix64(st) % (uint64_t)(i + 1));
                std::swap(perm[i], perm[k]);
            }
            pos = 0;
        }
        p = perm[pos++];
    }
    int *didx = nullptr;
    float *dl = nullptr;
    HIPCHK(hipMalloc((void **)&didx, sizeof(int) * picks.size()), "hipMalloc");
    hipError_t e = hipMalloc((void **)&dl, sizeof(float) * iters);
    if (e == hipSuccess)
        e = hipMemcpyAsync(didx, picks.data(), sizeof(int) * picks.size(), hipMemcpyHostToDevice,
                           s);
    if (e == hipSuccess) e = launch_fill(dl, (size_t)iters, 0.f, s);
    for (int it = 0; it < iters && e == hipSuccess; ++it) {
        e = launch_gather_bricks(bricks, didx + (size_t)it * batch, a->x0, batch, s);
        if (e == hipSuccess) e = ae_forward(a, a->x0, batch);
        if (e == hipSuccess) e = ae_backward(a, a->x0, batch, dl + it);
        a->steps++;
        const double c1 = 1.0 - std::pow((double)b1, (double)a->steps);
        const double c2 = 1.0 - std::pow((double)b2, (double)a->steps);
        const float lr_t = (float)(lr * std::sqrt(c2) / c1);
        if (e == hipSuccess)
            e = launch_adam(a->w, a->g, a->m, a->v, (int)kAeParams, lr_t, b1, b2, eps, s);
    }
    std::vector<float> lhost(iters);
    if (e == hipSuccess)
        e = hipMemcpyAsync(lhost.data(), dl, sizeof(float) * iters, hipMemcpyDeviceToHost, s);
    hipError_t es = hipStreamSynchronize(s);
    if (e == hipSuccess) e = es;
    (void)hipFree(didx);
    if (dl) (void)hipFree(dl);
    if (e != hipSuccess) return hip_fail(e, "autoencoder training");
    if (losses)
        for (int i = 0; i < iters; ++i)
            losses[i] = lhost[i] / (float)(kBrickCells * batch);  // mean squared error
    return RTMHIP_OK;
}

int rtmhip_ae_bricks(const float *field, int nxe, int nye, int nze, int x0, int y0, int z0,
                     int x1, int y1, int z1, void *bricks, float *scales)
{
    NEED_INIT();
    if (!field || !bricks || !scales) return fail(RTMHIP_ERR_INVALID, "NULL pointer");
    AeRegion r;
    Grid g;
    int rc = ae_check_region(nxe, nye, nze, x0, y0, z0, x1, y1, z1, &r, &g);
    if (rc) return rc;
    HIPCHK(launch_bricks(field, g, r, bricks, scales, g_state.stream), "bricks");
    return RTMHIP_OK;
}

int rtmhip_ae_encode(rtmhip_ae *a, const void *bricks, int n, void *codes)
{
    NEED_INIT();
    if (!a || !bricks || !codes) return fail(RTMHIP_ERR_INVALID, "NULL pointer");
    HIPCHK(ae_encode(a, bricks, n, codes), "encode");
    return RTMHIP_OK;
}

int rtmhip_ae_decode(rtmhip_ae *a, const void *codes, const float *scales, int n, float *field,
                     int nxe, int nye, int nze, int x0, int y0, int z0, int x1, int y1, int z1)
{
    NEED_INIT();
    if (!a || !codes || !scales || !field) return fail(RTMHIP_ERR_INVALID, "NULL pointer");
    AeRegion r;
    Grid g;
    int rc = ae_check_region(nxe, nye, nze, x0, y0, z0, x1, y1, z1, &r, &g);
    if (rc) return rc;
    if ((long long)n != (long long)r.nbx * r.nby * r.nbz)
        return fail(RTMHIP_ERR_INVALID, "%d codes for a region of %d bricks", n,
                    r.nbx * r.nby * r.nbz);
    HIPCHK(ae_decode(a, codes, scales, n, g, r, field), "decode");
    return RTMHIP_OK;
}

int rtmhip_snap_ae_sizes(const rtmhip_shot *shot, int k, int *bricks, size_t *code_bytes,
                         size_t *scale_bytes, size_t *work_bytes)
{
    if (!shot) return fail(RTMHIP_ERR_INVALID, "NULL pointer");
    if (shot->nt <= 0 || shot->nze <= 2 * shot->blen || shot->nxe <= 0 || shot->nye <= 0)
        return fail(RTMHIP_ERR_INVALID, "bad shot geometry");
    if (k < 1) return fail(RTMHIP_ERR_INVALID, "snapshot interval %d below 1", k);
    const long long nb = snap_region_bricks(shot);
    if (nb > 0x7fffffffLL) return fail(RTMHIP_ERR_INVALID, "imaging region of more than 2^31 bricks");
    const size_t nsnap = (size_t)((shot->nt - 1) / k + 1);
    if (bricks) *bricks = (int)nb;
    if (code_bytes) *code_bytes = nsnap * (size_t)nb * kCodes * 2;
    if (scale_bytes) *scale_bytes = nsnap * (size_t)nb * 4;
    if (work_bytes) *work_bytes = snap_ae_work_bytes(shot);
    return RTMHIP_OK;
}

int rtmhip_snap_ae_forward_steps(const rtmhip_shot *shot, const float *V2DT2, const float *coefs,
                                 const float *taper, const float *ricker, float **P, float **PP,
                                 rtmhip_ae *ae, void *codes, float *scales, void *work, int k,
                                 int it0, int nsteps)
{
    NEED_INIT();
    ShotCtx c;
    int rc;
    if ((rc = prepare_shot(shot, coefs, &c))) return rc;
    if ((rc = check_format(shot))) return rc;
    if (!V2DT2 || !ricker || !taper || !P || !PP || !*P || !*PP)
        return fail(RTMHIP_ERR_INVALID, "NULL pointer");
    if ((rc = check_snap_ae(shot, ae, codes, scales, work, k))) return rc;
    if (nsteps <= 0) return RTMHIP_OK;
    if (it0 < 0 || it0 + nsteps > shot->nt)
        return fail(RTMHIP_ERR_INVALID, "time range outside [0,%d)", shot->nt);
    const AeRegion r = snap_region(shot, shot->blen);
    const long long nb = snap_region_bricks(shot);
    return snap_forward_run(shot, c, V2DT2, taper, ricker, P, PP, k, it0, nsteps,
                            [&](const float *f, int j) {
                                HIPCHK(launch_bricks(f, c.g, r, work, scales + (size_t)j * nb,
                                                     g_state.stream),
                                       "snapshot bricks");
                                HIPCHK(ae_encode(ae, work, (int)nb,
                                                 static_cast<char *>(codes) +
                                                     (size_t)j * nb * kCodes * 2),
                                       "snapshot encode");
                                return RTMHIP_OK;
                            });
}

int rtmhip_snap_ae_backward_steps(const rtmhip_shot *shot, const float *V2DT2,
                                  const float *coefs, const float *taper, float **Pr, float **PPr,
                                  rtmhip_ae *ae, const void *codes, const float *scales,
                                  void *work, int k, const float *rcv_chunk, float *img,
                                  int it_hi, int nsteps)
{
    NEED_INIT();
    ShotCtx c;
    int rc;
    if ((rc = prepare_shot(shot, coefs, &c))) return rc;
    if ((rc = check_format(shot))) return rc;
    if (!V2DT2 || !taper || !Pr || !PPr || !*Pr || !*PPr || !rcv_chunk || !img)
        return fail(RTMHIP_ERR_INVALID, "NULL pointer");
    if ((rc = check_snap_ae(shot, ae, codes, scales, work, k))) return rc;
    if (nsteps <= 0) return RTMHIP_OK;
    if ((rc = check_chunk(shot, it_hi - nsteps + 1, it_hi))) return rc;
    const int lz = shot->nze - 2 * shot->blen;
    const AeRegion r = snap_region(shot, 0);
    const Grid cg = dense_grid(shot->nxe, shot->nye, lz);
    const long long nb = snap_region_bricks(shot);
    float *snap = static_cast<float *>(work);
    return snap_backward_run(shot, c, V2DT2, taper, Pr, PPr, k, rcv_chunk, it_hi, nsteps,
                             [&](const float *pr, int j) {
                                 HIPCHK(ae_decode(ae,
                                                  static_cast<const char *>(codes) +
                                                      (size_t)j * nb * kCodes * 2,
                                                  scales + (size_t)j * nb, (int)nb, cg, r, snap),
                                        "snapshot decode");
                                 HIPCHK(launch_snap_image(img, (const float *)snap, pr, c.g,
                                                          shot->blen, (float)k, g_state.stream),
                                        "snapshot image");
                                 return RTMHIP_OK;
                             });
}

int rtmhip_ae_gemm_test(const void *A, long long sam, long long sak, int a_half, const void *B,
                        long long sbk, long long sbn, int b_half, float *C, long long ldc,
                        const float *bias, long long M, int N, long long K, int act,
                        int accumulate)
{
    NEED_INIT();
    AeGemm p{};
    p.A = A;
    p.sam = sam;
    p.sak = sak;
    p.a_half = a_half;
    p.B = B;
    p.sbk = sbk;
    p.sbn = sbn;
    p.b_half = b_half;
    p.C = C;
    p.ldc = ldc;
    p.bias = bias;
    p.M = M;
    p.N = N;
    p.K = K;
    p.act = act;
    p.accumulate = accumulate;
    p.kchunk = 64;
    HIPCHK(launch_ae_gemm(p, g_state.stream), "gemm");
    return RTMHIP_OK;
}

}  // extern "C"
