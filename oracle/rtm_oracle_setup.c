/*
 * rtm_oracle_setup.c -- host-side set-up of the RTM oracle (TEST INFRASTRUCTURE).
 *
 * FP32 functions restate the reference's C++ expressions with the same type
 * promotions (float vs double intermediates) so that the tables they produce
 * are bit-identical to the reference program's on the same libm.
 * Compiled with -ffp-contract=off and without -ffast-math.
 */
#include <math.h>
#include <string.h>
#include "rtm_oracle.h"

/* rtm-3d/src/misc/Misc.cpp:106-137 (orders other than 2,4,6,8). */
static void makeo2_f32(float *coef, int order)
{
    float h_beta, alpha1, alpha2, central = 0.0f, coef_filt, arg, coef_wind;
    float alpha = .54f, beta = 6.f;
    int msign = -1, ix;
    h_beta = (float)(0.5 * beta);
    alpha1 = (float)(2. * alpha - 1.0);
    alpha2 = (float)(2. * (1.0 - alpha));
    for (ix = 1; ix <= order / 2; ix++) {
        msign = -msign;
        coef_filt = (float)((2. * msign) / (ix * ix));
        arg = (float)(M_PI * ix / (2. * (order / 2 + 2)));
        coef_wind = powf(alpha1 + alpha2 * cosf(arg) * cosf(arg), h_beta);
        coef[order / 2 + ix] = coef_filt * coef_wind;
        central = central + coef[order / 2 + ix];
        coef[order / 2 - ix] = coef[order / 2 + ix];
    }
    coef[order / 2] = (float)(-2. * central);
}

/* rtm-3d/src/misc/Misc.cpp:139-176: double literals stored as float. */
void orc_laplacian_taps_f32(int order, float *c)
{
    switch (order) {
    case 2: c[0] = 1.f; c[1] = -2.f; c[2] = 1.f; break;
    case 4:
        c[0] = (float)(-1. / 12.); c[1] = (float)(4. / 3.); c[2] = (float)(-5. / 2.);
        c[3] = (float)(4. / 3.); c[4] = (float)(-1. / 12.); break;
    case 6:
        c[0] = (float)(1. / 90.); c[1] = (float)(-3. / 20.); c[2] = (float)(3. / 2.);
        c[3] = (float)(-49. / 18.); c[4] = (float)(3. / 2.); c[5] = (float)(-3. / 20.);
        c[6] = (float)(1. / 90.); break;
    case 8:
        c[0] = (float)(-1. / 560.); c[1] = (float)(8. / 315.); c[2] = (float)(-1. / 5.);
        c[3] = (float)(8. / 5.); c[4] = (float)(-205. / 72.); c[5] = (float)(8. / 5.);
        c[6] = (float)(-1. / 5.); c[7] = (float)(8. / 315.); c[8] = (float)(-1. / 560.); break;
    default: makeo2_f32(c, order);
    }
}

/* rtm-3d/include/RTMGrid.hpp:433-446: devinv = (1.0/d)*(1.0/d) in double,
 * stored as float; tap = devinv * coef in float. Layout [dim][k]. */
void orc_stencil_coefs_f32(int order, float dx, float dy, float dz, float *out)
{
    float taps[64];
    float d[3] = {dx, dy, dz};
    orc_laplacian_taps_f32(order, taps);
    for (int k = 0; k < 3; k++) {
        float devinv = (float)((1.0 / (double)d[k]) * (1.0 / (double)d[k]));
        for (int io = 0; io <= order; io++) out[k * (order + 1) + io] = devinv * taps[io];
    }
}

/* rtm-3d/include/RTMSeismic.hpp:135-144 and :191-198.
 * t = (float)((double)((float)it*dt) - 1.0/fpeak); x = (float)(pi*fpeak*t);
 * s = (float)((double)expf(-x*x) * (1.0 - 2.0*x*x)). */
void orc_ricker_f32(int nt, float dt, float fpeak, float *out)
{
    for (int it = 0; it < nt; it++) {
        float t = (float)((double)((float)it * dt) - 1.0 / (double)fpeak);
        float x = (float)(M_PI * (double)fpeak * (double)t);
        float xx = x * x;
        out[it] = (float)((double)expf(-xx) * (1.0 - 2.0 * (double)xx));
    }
}

/* rtm-3d/include/RTMSeismic.hpp:294-305: dfrac in float from float log/sqrt,
 * g[i] = (float)exp(-pow((double)(dfrac*i), 2)). */
void orc_taper_table_f32(int blen, float factor, float *out)
{
    float dfrac = (float)((double)sqrtf(-logf(factor)) / (1. * blen));
    for (int i = 0; i < blen; i++) {
        double a = (double)(dfrac * (float)i);
        out[i] = (float)exp(-pow(a, 2.0));
    }
}

/* glibc random_r TYPE_3 (degree 31, separation 3): seeded by the
 * Park-Miller minimal standard generator, 310 discarded draws, then
 * r[f] += r[b]; output r[f] >> 1. */
void orc_glibc_rand_seed(orc_rand_state *st, unsigned seed)
{
    int32_t word = (int32_t)(seed == 0 ? 1 : seed);
    memset(st, 0, sizeof(*st));
    st->tbl[0] = word;
    for (int i = 1; i < 31; i++) {
        int32_t hi = word / 127773, lo = word % 127773;
        word = 16807 * lo - 2836 * hi;
        if (word < 0) word += 2147483647;
        st->tbl[i] = word;
    }
    st->f = 3;
    st->b = 0;
    for (int i = 0; i < 310; i++) (void)orc_glibc_rand_next(st);
}

int orc_glibc_rand_next(orc_rand_state *st)
{
    uint32_t v = (uint32_t)st->tbl[st->f] + (uint32_t)st->tbl[st->b];
    st->tbl[st->f] = (int32_t)v;
    st->f = (st->f + 1) % 31;
    st->b = (st->b + 1) % 31;
    return (int)(v >> 1);
}

/* rtm-3d/src/misc/Misc.cpp:91-102 (reached through randbetween, :35-38). */
float orc_randlinearval(float v, int k, int nkb, orc_rand_state *st)
{
    float vrand = v;
    float dist = (float)(k * 1.0);
    float range = (float)(nkb * 1.0);
    float perc = (dist / range) * 100;
    float seed = (float)((orc_glibc_rand_next(st) % 100) * 1.0);
    if (seed < perc && perc > 70.)
        vrand = (float)((double)v * (1.0 / (double)perc) + (double)v * ((double)seed / 100.0));
    return vrand;
}

#define IDX(x, y, z) (((size_t)(x) * (size_t)nye + (size_t)(y)) * (size_t)nze + (size_t)(z))

/* rtm-3d/src/grid/RTMGrid.cpp:47-86 (interior placement). */
void orc_extend_borders_f32(const float *v, int nx, int ny, int nz, int blen, float *ve)
{
    int nxe = nx + 2 * blen, nye = ny + 2 * blen, nze = nz + 2 * blen;
    memset(ve, 0, sizeof(float) * (size_t)nxe * nye * nze);
    for (int x = 0; x < nx; x++)
        for (int y = 0; y < ny; y++)
            memcpy(ve + IDX(x + blen, y + blen, blen), v + ((size_t)x * ny + y) * nz,
                   sizeof(float) * (size_t)nz);
}

/* rtm-3d/src/grid/RTMGrid.cpp:371-434: clamp every index to the interior. */
void orc_borders_abc_f32(float *ve, int nxe, int nye, int nze, int blen)
{
    int lx = nxe - 2 * blen, ly = nye - 2 * blen, lz = nze - 2 * blen;
    for (int x = 0; x < nxe; x++) {
        int kx = x <= blen ? blen : (x >= lx + blen ? lx + blen - 1 : x);
        for (int y = 0; y < nye; y++) {
            int ky = y <= blen ? blen : (y >= ly + blen ? ly + blen - 1 : y);
            for (int z = 0; z < nze; z++) {
                int kz = z <= blen ? blen : (z >= lz + blen ? lz + blen - 1 : z);
                ve[IDX(x, y, z)] = ve[IDX(kx, ky, kz)];
            }
        }
    }
}

/* rtm-3d/src/grid/RTMGrid.cpp:437-635, executed in the sequential loop order
 * that one OpenMP thread produces (the reference is only deterministic then). */
void orc_borders_rbc_f32(float *ve, int nxe, int nye, int nze, int blen, orc_rand_state *st)
{
    const int nx = nxe, ny = nye, nz = nze, b = blen;
#define RB(v, k) orc_randlinearval((v), (int)(k), b, st)
    for (int ix = b; ix < nx - b; ix++)
        for (int iy = b; iy < ny - b; iy++)
            for (int iz = 0; iz < b; iz++) {
                int dist = b - iz;
                ve[IDX(ix, iy, iz)] = RB(ve[IDX(ix, iy, b)], dist);
                ve[IDX(ix, iy, nz - 1 - iz)] = RB(ve[IDX(ix, iy, nz - b - 1)], dist);
            }
    for (int ix = b; ix < nx - b; ix++)
        for (int iz = b; iz < nz - b; iz++)
            for (int iy = 0; iy < b; iy++) {
                ve[IDX(ix, iy, iz)] = RB(ve[IDX(ix, b, iz)], b - iy);
                ve[IDX(ix, ny - b + iy, iz)] = RB(ve[IDX(ix, ny - b - 1, iz)], iy);
            }
    for (int iy = b; iy < ny - b; iy++)
        for (int iz = b; iz < nz - b; iz++)
            for (int ix = 0; ix < b; ix++) {
                ve[IDX(ix, iy, iz)] = RB(ve[IDX(b, iy, iz)], b - ix);
                ve[IDX(nx - b + ix, iy, iz)] = RB(ve[IDX(nx - b - 1, iy, iz)], ix);
            }
    for (int ix = 0; ix < b; ix++)
        for (int iz = 0; iz < b; iz++)
            for (int iy = 0; iy < ny; iy++) {
                int ky = iy < b ? b : (iy >= ny - b ? ny - b - 1 : iy);
                /* four corner columns; y factor first, then x, then z */
                const int xs[4] = {b, nx - b - 1, b, nx - b - 1};
                const int zs[4] = {b, b, nz - b - 1, nz - b - 1};
                const int ox[4] = {ix, nx - b + ix, ix, nx - b + ix};
                const int oz[4] = {iz, iz, nz - b + iz, nz - b + iz};
                const int kx[4] = {b - ix, ix, b - ix, ix};
                const int kz[4] = {b - iz, b - iz, iz, iz};
                for (int c = 0; c < 4; c++) {
                    float v = ve[IDX(xs[c], ky, zs[c])];
                    if (iy < b) v = RB(v, b - iy);
                    else if (iy >= ny - b) v = RB(v, iy - (ny - b));
                    v = RB(v, kx[c]);
                    v = RB(v, kz[c]);
                    ve[IDX(ox[c], iy, oz[c])] = v;
                }
            }
    for (int iy = 0; iy < b; iy++)
        for (int iz = 0; iz < b; iz++)
            for (int ix = b; ix < nx - b; ix++) {
                float v;
                v = RB(RB(ve[IDX(ix, b, b)], b - iy), b - iz);
                ve[IDX(ix, iy, iz)] = v;
                v = RB(RB(ve[IDX(ix, ny - b - 1, b)], iy), b - iz);
                ve[IDX(ix, ny - b + iy, iz)] = v;
                v = RB(RB(ve[IDX(ix, b, nz - b - 1)], b - iy), iz);
                ve[IDX(ix, iy, nz - b + iz)] = v;
                v = RB(RB(ve[IDX(ix, ny - b - 1, nz - b - 1)], iy), iz);
                ve[IDX(ix, ny - b + iy, nz - b + iz)] = v;
            }
    for (int ix = 0; ix < b; ix++)
        for (int iy = 0; iy < b; iy++)
            for (int iz = b; iz < nz - b; iz++) {
                float v;
                v = RB(RB(ve[IDX(b, b, iz)], b - ix), b - iy);
                ve[IDX(ix, iy, iz)] = v;
                v = RB(RB(ve[IDX(b, ny - b - 1, iz)], b - ix), iy);
                ve[IDX(ix, ny - b + iy, iz)] = v;
                v = RB(RB(ve[IDX(nx - b - 1, b, iz)], ix), b - iy);
                ve[IDX(nx - b + ix, iy, iz)] = v;
                v = RB(RB(ve[IDX(nx - b - 1, ny - b - 1, iz)], ix), iy);
                ve[IDX(nx - b + ix, ny - b + iy, iz)] = v;
            }
#undef RB
}

/* rtm-3d/src/grid/RTMGrid.cpp:638-696: ABC, then RBC, then the top slab is a
 * clamped copy of the z = blen plane. */
void orc_borders_hbc_f32(float *ve, int nxe, int nye, int nze, int blen, orc_rand_state *st)
{
    int lx = nxe - 2 * blen, ly = nye - 2 * blen;
    orc_borders_abc_f32(ve, nxe, nye, nze, blen);
    orc_borders_rbc_f32(ve, nxe, nye, nze, blen, st);
    for (int x = 0; x < nxe; x++) {
        int kx = x <= blen ? blen : (x >= lx + blen ? lx + blen - 1 : x);
        for (int y = 0; y < nye; y++) {
            int ky = y <= blen ? blen : (y >= ly + blen ? ly + blen - 1 : y);
            for (int z = 0; z < blen; z++) ve[IDX(x, y, z)] = ve[IDX(kx, ky, blen)];
        }
    }
}

/* rtm-3d/include/Grid.hpp:212-235 via RTMController.cpp:110-111. */
void orc_v2dt2_f32(float *v, size_t n, float dt)
{
    float dt2 = dt * dt;
    for (size_t i = 0; i < n; i++) {
        v[i] *= v[i];
        v[i] *= dt2;
    }
}

/* Frequency tables of RTMFreqDomainKernel::initKernel
 * (rtm-3d/src/kernel/RTMFiniteDifferencesKernel_FreqDomain.cpp:125-152),
 * rtmInitWList (:17-33) and rtmInitFreqKernels (:35-52), with the
 * reference's float members (freq_min, freq_max, df) and its float/double
 * promotions: df = 1/(double(nt)*dt) stored as float; nw = size_t of the
 * float (fmax - fmin)/df; w[iw] = float(2*double(pi_f)*((float(iw)*df) +
 * fmin)*double(dt)); kre/kim[iw][it] = cosf/sinf(float(double(w)*it));
 * w2 = w*w in float.  Returns nw; the tables are written when nw <= cap. */
int orc_freq_tables(int nt, float dt, float fmin, float fmax, float *w2, float *kre, float *kim,
                    int cap)
{
    const float freq_max = 1.0 * fmax;
    const float freq_min = 1.0 * fmin;
    const float df = 1. / ((nt * 1.0) * dt);
    const float nf = ((freq_max - freq_min) / df);
    const int nw = (int)(size_t)nf;
    if (nw > cap) return nw;
    const float pi = acosf(-1.);
    for (int iw = 0; iw < nw; iw++) {
        const float lw = (float)iw * 1.0;
        const float w = (float)(2. * pi * ((lw * df) + freq_min) * dt);
        for (int it = 0; it < nt; it++) {
            kre[(size_t)iw * nt + it] = cosf(w * (it * 1.0));
            kim[(size_t)iw * nt + it] = sinf(w * (it * 1.0));
        }
        w2[iw] = w * w;
    }
    return nw;
}

/* ---------------- double-precision set-up (ideal values) -------------- */
void orc_stencil_coefs_f64(int order, double dx, double dy, double dz, double *out)
{
    double taps[64];
    double d[3] = {dx, dy, dz};
    switch (order) {
    case 2: taps[0] = 1; taps[1] = -2; taps[2] = 1; break;
    case 4: taps[0] = -1. / 12; taps[1] = 4. / 3; taps[2] = -5. / 2; taps[3] = 4. / 3;
            taps[4] = -1. / 12; break;
    case 6: taps[0] = 1. / 90; taps[1] = -3. / 20; taps[2] = 3. / 2; taps[3] = -49. / 18;
            taps[4] = 3. / 2; taps[5] = -3. / 20; taps[6] = 1. / 90; break;
    default:
        taps[0] = -1. / 560; taps[1] = 8. / 315; taps[2] = -1. / 5; taps[3] = 8. / 5;
        taps[4] = -205. / 72; taps[5] = 8. / 5; taps[6] = -1. / 5; taps[7] = 8. / 315;
        taps[8] = -1. / 560; break;
    }
    for (int k = 0; k < 3; k++)
        for (int io = 0; io <= order; io++) out[k * (order + 1) + io] = taps[io] / (d[k] * d[k]);
}

void orc_ricker_f64(int nt, double dt, double fpeak, double *out)
{
    for (int it = 0; it < nt; it++) {
        double x = M_PI * fpeak * (it * dt - 1.0 / fpeak);
        out[it] = exp(-x * x) * (1.0 - 2.0 * x * x);
    }
}

void orc_taper_table_f64(int blen, double factor, double *out)
{
    double dfrac = sqrt(-log(factor)) / blen;
    for (int i = 0; i < blen; i++) out[i] = exp(-(dfrac * i) * (dfrac * i));
}

void orc_v2dt2_f64(const float *v, size_t n, double dt, double *out)
{
    for (size_t i = 0; i < n; i++) out[i] = (double)v[i] * (double)v[i] * dt * dt;
}

/* ---------------- 16-bit storage conversions ----------------------------
 * IEEE 754 binary16 and bfloat16 (the top half of binary32), both rounded
 * to nearest, ties to even, with subnormals kept and overflow to infinity.
 * Written from the format definitions; pinned in the tests against numpy's
 * float16 casts (exhaustively for the widening direction) and torch's
 * bfloat16 casts. */
uint16_t orc_f32_to_f16(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    const uint32_t sign = (u >> 16) & 0x8000u, a = u & 0x7fffffffu;
    if (a > 0x7f800000u) return (uint16_t)(sign | 0x7e00u | ((a >> 13) & 0x3ffu)); /* NaN */
    if (a >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u); /* >= 65520 rounds to inf */
    if (a >= 0x38800000u) {                                  /* normal: >= 2^-14 */
        const uint32_t m = a - 0x38000000u;                  /* exponent rebias 127 -> 15 */
        return (uint16_t)(sign | ((m + 0xfffu + ((m >> 13) & 1u)) >> 13));
    }
    if (a < 0x33000000u) return (uint16_t)sign; /* below half the smallest subnormal */
    const int e = (int)(a >> 23);               /* 102..112 */
    const uint32_t mant = (a & 0x7fffffu) | 0x800000u;
    const int shift = 126 - e;                  /* value / 2^-24 = mant >> shift */
    uint32_t q = mant >> shift;
    const uint32_t rem = mant & ((1u << shift) - 1u), half = 1u << (shift - 1);
    if (rem > half || (rem == half && (q & 1u))) q++;
    return (uint16_t)(sign | q);
}

float orc_f16_to_f32(uint16_t h)
{
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu, u;
    if (e == 0x1fu) {
        u = sign | 0x7f800000u | (m << 13);
    } else if (e) {
        u = sign | ((e + 112u) << 23) | (m << 13);
    } else if (!m) {
        u = sign;
    } else {
        int s = 0;
        while (!(m & 0x400u)) { m <<= 1; s++; }
        u = sign | ((uint32_t)(113 - s) << 23) | ((m & 0x3ffu) << 13);
    }
    float f;
    memcpy(&f, &u, 4);
    return f;
}

uint16_t orc_f32_to_bf16(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u); /* quiet NaN */
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

float orc_bf16_to_f32(uint16_t b)
{
    const uint32_t u = (uint32_t)b << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

void orc_to_f16(const float *in, uint16_t *out, size_t n)
{
    for (size_t i = 0; i < n; i++) out[i] = orc_f32_to_f16(in[i]);
}
void orc_from_f16(const uint16_t *in, float *out, size_t n)
{
    for (size_t i = 0; i < n; i++) out[i] = orc_f16_to_f32(in[i]);
}
void orc_to_bf16(const float *in, uint16_t *out, size_t n)
{
    for (size_t i = 0; i < n; i++) out[i] = orc_f32_to_bf16(in[i]);
}
void orc_from_bf16(const uint16_t *in, float *out, size_t n)
{
    for (size_t i = 0; i < n; i++) out[i] = orc_bf16_to_f32(in[i]);
}
