/*
 * rtm_oracle_ops.c -- per-step operators and time loops of the RTM oracle.
 * TEST INFRASTRUCTURE ONLY (see rtm_oracle.h).
 *
 * Compiled four times: -DORC_REAL=float -DORC_SFX=f32, -DORC_REAL=double
 * -DORC_SFX=f64, and the 16-bit storage builds -DORC_REAL=float with
 * -DORC_SFX=f16 -DORC_STORE_F16 or -DORC_SFX=bf16 -DORC_STORE_BF16.  The
 * float build restates the reference CPU platform
 * (rtm-3d/src/platform/host/RTMCPUPlatform.cpp) operation for operation;
 * -ffp-contract=off keeps every multiply and add separately rounded.
 *
 * Wavefields (P, PP, the UPB layer) have storage type W.  In the 16-bit
 * builds every operator computes in float exactly as the float build and
 * every wavefield write rounds to the storage format (nearest, ties to
 * even): the reference's operator sequence with 16-bit grids.  Velocity,
 * taps, traces and the image stay float.
 */
#include <stdlib.h>
#include <string.h>
#include "rtm_oracle.h"

#ifndef ORC_REAL
#error "compile with -DORC_REAL=float|double -DORC_SFX=f32|f64"
#endif
typedef ORC_REAL R;
#if defined(ORC_STORE_F16)
typedef uint16_t W;
#define LD(v) orc_f16_to_f32(v)
#define ST(x) orc_f32_to_f16(x)
#elif defined(ORC_STORE_BF16)
typedef uint16_t W;
#define LD(v) orc_bf16_to_f32(v)
#define ST(x) orc_f32_to_bf16(x)
#else
typedef R W;
#define LD(v) (v)
#define ST(x) (x)
#endif
#define CAT2(a, b) a##_##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, ORC_SFX)

#define IDX(x, y, z) (((size_t)(x) * (size_t)nye + (size_t)(y)) * (size_t)nze + (size_t)(z))

/* rtm-3d/src/platform/host/RTMCPUPlatform.cpp:325-395 (rtmStep3D).
 * acm_d accumulated from 0 in ascending tap order; lapl = (acmx+acmy)+acmz;
 * PP <- ((2*p) - pp) + v2dt2*lapl, in place.  Returns -1 on NaN
 * ("Numerical Dispersion"), 0 otherwise. */
int FN(orc_step)(const W *P, W *PP, const R *V2, const R *coefs, int nxe, int nye, int nze,
                 int order)
{
    const int hf = order / 2;
    const R *cx = coefs, *cy = coefs + (order + 1), *cz = coefs + 2 * (order + 1);
    const size_t sx = (size_t)nye * nze, sy = (size_t)nze;
    int bad = 0;
#pragma omp parallel for collapse(2) schedule(static) reduction(| : bad)
    for (int ix = hf; ix < nxe - hf; ix++)
        for (int iy = hf; iy < nye - hf; iy++) {
            size_t o = IDX(ix, iy, hf);
            for (int iz = hf; iz < nze - hf; iz++, o++) {
                R ax = 0.0, ay = 0.0, az = 0.0;
                for (int io = 0; io <= order; io++) {
                    ptrdiff_t d = io - hf;
                    ax += LD(P[o + d * (ptrdiff_t)sx]) * cx[io];
                    ay += LD(P[o + d * (ptrdiff_t)sy]) * cy[io];
                    az += LD(P[o + d]) * cz[io];
                }
                R lapl = ax + ay + az;
                R p = LD(P[o]), pp = LD(PP[o]);
                R npp = 2 * p - pp + V2[o] * lapl;
                PP[o] = ST(npp);
                if (npp != npp) bad = 1;
            }
        }
    return bad ? -1 : 0;
}

/* rtm-3d/src/platform/host/RTMCPUPlatform.cpp:710-920 (taper3D).  The loop
 * nests run in the reference's order; the front-right z-edge reads the
 * already face-tapered cell (blen, ny-blen+iy, iz) instead of its own. */
void FN(orc_taper_all)(W *P, const R *g, int nxe, int nye, int nze, int blen)
{
    const int nx = nxe, ny = nye, nz = nze, b = blen;
#pragma omp parallel for collapse(2)
    for (int ix = b; ix < nx - b; ix++)
        for (int iy = b; iy < ny - b; iy++)
            for (int iz = 0; iz < b; iz++) {
                P[IDX(ix, iy, iz)] = ST(LD(P[IDX(ix, iy, iz)]) * g[b - 1 - iz]);
                P[IDX(ix, iy, nz - b + iz)] = ST(LD(P[IDX(ix, iy, nz - b + iz)]) * g[iz]);
            }
#pragma omp parallel for collapse(2)
    for (int ix = b; ix < nx - b; ix++)
        for (int iz = b; iz < nz - b; iz++)
            for (int iy = 0; iy < b; iy++) {
                P[IDX(ix, iy, iz)] = ST(LD(P[IDX(ix, iy, iz)]) * g[b - 1 - iy]);
                P[IDX(ix, ny - b + iy, iz)] = ST(LD(P[IDX(ix, ny - b + iy, iz)]) * g[iy]);
            }
#pragma omp parallel for collapse(2)
    for (int iy = b; iy < ny - b; iy++)
        for (int iz = b; iz < nz - b; iz++)
            for (int ix = 0; ix < b; ix++) {
                P[IDX(ix, iy, iz)] = ST(LD(P[IDX(ix, iy, iz)]) * g[b - 1 - ix]);
                P[IDX(nx - b + ix, iy, iz)] = ST(LD(P[IDX(nx - b + ix, iy, iz)]) * g[ix]);
            }
#pragma omp parallel for collapse(2)
    for (int ix = 0; ix < b; ix++)
        for (int iz = 0; iz < b; iz++)
            for (int iy = 0; iy < ny; iy++) {
                const int has_y = (iy < b) || (iy >= ny - b);
                const R gy = iy < b ? g[b - 1 - iy] : (iy >= ny - b ? g[iy - (ny - b)] : (R)1);
                R v;
                v = LD(P[IDX(ix, iy, iz)]);
                if (has_y) v *= gy;
                v *= g[b - 1 - ix]; v *= g[b - 1 - iz];
                P[IDX(ix, iy, iz)] = ST(v);
                v = LD(P[IDX(nx - b + ix, iy, iz)]);
                if (has_y) v *= gy;
                v *= g[ix]; v *= g[b - 1 - iz];
                P[IDX(nx - b + ix, iy, iz)] = ST(v);
                v = LD(P[IDX(ix, iy, nz - b + iz)]);
                if (has_y) v *= gy;
                v *= g[b - 1 - ix]; v *= g[iz];
                P[IDX(ix, iy, nz - b + iz)] = ST(v);
                v = LD(P[IDX(nx - b + ix, iy, nz - b + iz)]);
                if (has_y) v *= gy;
                v *= g[ix]; v *= g[iz];
                P[IDX(nx - b + ix, iy, nz - b + iz)] = ST(v);
            }
#pragma omp parallel for collapse(2)
    for (int iy = 0; iy < b; iy++)
        for (int iz = 0; iz < b; iz++)
            for (int ix = b; ix < nx - b; ix++) {
                R v;
                v = LD(P[IDX(ix, iy, iz)]); v *= g[b - 1 - iy]; v *= g[b - 1 - iz];
                P[IDX(ix, iy, iz)] = ST(v);
                v = LD(P[IDX(ix, ny - b + iy, iz)]); v *= g[iy]; v *= g[b - 1 - iz];
                P[IDX(ix, ny - b + iy, iz)] = ST(v);
                v = LD(P[IDX(ix, iy, nz - b + iz)]); v *= g[b - 1 - iy]; v *= g[iz];
                P[IDX(ix, iy, nz - b + iz)] = ST(v);
                v = LD(P[IDX(ix, ny - b + iy, nz - b + iz)]); v *= g[iy]; v *= g[iz];
                P[IDX(ix, ny - b + iy, nz - b + iz)] = ST(v);
            }
#pragma omp parallel for collapse(2)
    for (int ix = 0; ix < b; ix++)
        for (int iy = 0; iy < b; iy++)
            for (int iz = b; iz < nz - b; iz++) {
                R v;
                v = LD(P[IDX(ix, iy, iz)]); v *= g[b - 1 - ix]; v *= g[b - 1 - iy];
                P[IDX(ix, iy, iz)] = ST(v);
                v = LD(P[IDX(b, ny - b + iy, iz)]); v *= g[b - 1 - ix]; v *= g[iy];
                P[IDX(ix, ny - b + iy, iz)] = ST(v);
                v = LD(P[IDX(nx - b + ix, iy, iz)]); v *= g[ix]; v *= g[b - 1 - iy];
                P[IDX(nx - b + ix, iy, iz)] = ST(v);
                v = LD(P[IDX(nx - b + ix, ny - b + iy, iz)]); v *= g[ix]; v *= g[iy];
                P[IDX(nx - b + ix, ny - b + iy, iz)] = ST(v);
            }
}

/* rtm-3d/src/platform/host/RTMCPUPlatform.cpp:566-657 (taperUpper3D). */
void FN(orc_taper_upper)(W *P, const R *g, int nxe, int nye, int nze, int blen)
{
    const int nx = nxe, ny = nye, b = blen;
#pragma omp parallel for collapse(2)
    for (int ix = b; ix < nx - b; ix++)
        for (int iy = b; iy < ny - b; iy++)
            for (int iz = 0; iz < b; iz++) P[IDX(ix, iy, iz)] = ST(LD(P[IDX(ix, iy, iz)]) * g[b - 1 - iz]);
#pragma omp parallel for collapse(2)
    for (int ix = 0; ix < b; ix++)
        for (int iz = 0; iz < b; iz++)
            for (int iy = 0; iy < ny; iy++) {
                const int has_y = (iy < b) || (iy >= ny - b);
                const R gy = iy < b ? g[b - 1 - iy] : (iy >= ny - b ? g[iy - (ny - b)] : (R)1);
                R v;
                v = LD(P[IDX(ix, iy, iz)]);
                if (has_y) v *= gy;
                v *= g[b - 1 - ix]; v *= g[b - 1 - iz];
                P[IDX(ix, iy, iz)] = ST(v);
                v = LD(P[IDX(nx - b + ix, iy, iz)]);
                if (has_y) v *= gy;
                v *= g[ix]; v *= g[b - 1 - iz];
                P[IDX(nx - b + ix, iy, iz)] = ST(v);
            }
#pragma omp parallel for collapse(2)
    for (int iy = 0; iy < b; iy++)
        for (int iz = 0; iz < b; iz++)
            for (int ix = b; ix < nx - b; ix++) {
                R v;
                v = LD(P[IDX(ix, iy, iz)]); v *= g[b - 1 - iy]; v *= g[b - 1 - iz];
                P[IDX(ix, iy, iz)] = ST(v);
                v = LD(P[IDX(ix, ny - b + iy, iz)]); v *= g[iy]; v *= g[b - 1 - iz];
                P[IDX(ix, ny - b + iy, iz)] = ST(v);
            }
}

/* rtm-3d/src/platform/host/RTMCPUPlatform.cpp:103-113 and
 * rtm-3d/src/grid/RTMGrid.cpp:36-40 (coordinates are extended ones). */
void FN(orc_apply_source)(W *PP, int nxe, int nye, int nze, int sxe, int sye, int sze, R val)
{
    if (sxe < 0 || sxe >= nxe || sye < 0 || sye >= nye || sze < 0 || sze >= nze) return;
    PP[IDX(sxe, sye, sze)] = ST(LD(PP[IDX(sxe, sye, sze)]) + val);
}

/* rtm-3d/src/platform/host/RTMCPUPlatform.cpp:114-151.  rcv is one ntstep
 * chunk laid out [nrx][nry][ntstep]; restore SETs the wavefield value (the
 * CPU semantics, not the CUDA accumulate). */
void FN(orc_receivers)(W *PP, R *rcv, int nxe, int nye, int nze, int blen, const orc_rcv_geom *g,
                       int ntstep, int lt, int restore)
{
    for (int ix = 0; ix < g->count_x; ix++)
        for (int iy = 0; iy < g->count_y; iy++) {
            int rx = g->start_x + ix * g->dist_x + blen;
            int ry = g->start_y + iy * g->dist_y + blen;
            int rz = g->depth_z + blen;
            if (rx < 0 || rx >= nxe || ry < 0 || ry >= nye || rz < 0 || rz >= nze) continue;
            size_t ro = ((size_t)ix * g->count_y + iy) * (size_t)ntstep + lt;
            if (restore) PP[IDX(rx, ry, rz)] = ST(rcv[ro]);
            else rcv[ro] = LD(PP[IDX(rx, ry, rz)]);
        }
}

/* rtm-3d/src/platform/host/RTMCPUPlatform.cpp:153-189.  upb is
 * [ntstep][nxe][nye][order/2]; the layer is z in [blen-order/2, blen). */
void FN(orc_upb)(W *PP, W *upb, int nxe, int nye, int nze, int blen, int order, int lt,
                 int restore)
{
    const int hf = order / 2;
#pragma omp parallel for collapse(2)
    for (int ix = 0; ix < nxe; ix++)
        for (int iy = 0; iy < nye; iy++)
            for (int k = 0; k < hf; k++) {
                size_t uo = (((size_t)lt * nxe + ix) * (size_t)nye + iy) * hf + k;
                size_t po = IDX(ix, iy, blen - hf + k);
                if (restore) PP[po] = upb[uo];
                else upb[uo] = PP[po];
            }
}

/* rtm-3d/src/platform/host/RTMCPUPlatform.cpp:513-540: all x, y and
 * z in [blen, nze-blen); img <- img + PS*PR. */
void FN(orc_image)(R *img, const W *PS, const W *PR, int nxe, int nye, int nze, int blen)
{
#pragma omp parallel for collapse(2)
    for (int ix = 0; ix < nxe; ix++)
        for (int iy = 0; iy < nye; iy++)
            for (int iz = blen; iz < nze - blen; iz++) {
                size_t o = IDX(ix, iy, iz);
                img[o] = img[o] + LD(PS[o]) * LD(PR[o]);
            }
}

/* rtm-3d/src/grid/RTMGrid.cpp:226-300 (filter, order 2) followed by
 * removeBorders(blen) (:129-166).  fcoefs is the [3][3] plane built like the
 * stencil plane from {1,-2,1}; out is [nx][ny][nz]. */
void FN(orc_image_finish)(const R *img, R *out, int nxe, int nye, int nze, int blen,
                          const R *fcoefs)
{
    const int nx = nxe - 2 * blen, ny = nye - 2 * blen, nz = nze - 2 * blen;
    const size_t sx = (size_t)nye * nze, sy = (size_t)nze;
#pragma omp parallel for collapse(2)
    for (int x = 0; x < nx; x++)
        for (int y = 0; y < ny; y++)
            for (int z = 0; z < nz; z++) {
                size_t o = IDX(x + blen, y + blen, z + blen);
                R ax = 0.0, ay = 0.0, az = 0.0;
                for (int io = 0; io < 3; io++) {
                    ptrdiff_t d = io - 1;
                    ax += img[o + d * (ptrdiff_t)sx] * fcoefs[io];
                    ay += img[o + d * (ptrdiff_t)sy] * fcoefs[3 + io];
                    az += img[o + d] * fcoefs[6 + io];
                }
                out[((size_t)x * ny + y) * nz + z] = ax + ay + az;
            }
}

/* rtmUpdateFreqContributions then rtmFreqDomainImageCondition
 * (rtm-3d/src/platform/host/RTMCPUPlatform.cpp:20-96) for the frequencies
 * iw = iw0 + lw (lw < nwstep, iw < nw) at time index st: on the interior
 * (x < nx, y < ny, all nz, extended offset +blen), in frequency order,
 *   sRe += kre*ps; sIm += kim*ps; rRe += kre*pr; rIm += kim*pr;
 *   img += w2 * (sRe*rRe - sIm*rIm).
 * Accumulators are [nwstep][nx][ny][nz].  The update of one frequency over
 * the whole grid precedes its imaging in the reference; both are per cell,
 * so doing them cell by cell gives the same values. */
void FN(orc_freq_image)(R *img, const W *PS, const W *PR, R *sre, R *sim, R *rre, R *rim,
                        const R *kre, const R *kim, const R *w2, int nxe, int nye, int nze,
                        int blen, int nwstep, int iw0, int nw, int nt, int st)
{
    const int nx = nxe - 2 * blen, ny = nye - 2 * blen, nz = nze - 2 * blen;
    const size_t vol = (size_t)nx * ny * nz;
    for (int lw = 0, iw = iw0; lw < nwstep && iw < nw; lw++, iw++) {
        const R kr = kre[(size_t)iw * nt + st], ki = kim[(size_t)iw * nt + st], w = w2[iw];
        R *a = sre + lw * vol, *b = sim + lw * vol, *c = rre + lw * vol, *d = rim + lw * vol;
#pragma omp parallel for collapse(2) schedule(static)
        for (int ix = 0; ix < nx; ix++)
            for (int iy = 0; iy < ny; iy++) {
                size_t f = ((size_t)ix * ny + iy) * nz;
                size_t o = IDX(ix + blen, iy + blen, blen);
                for (int iz = 0; iz < nz; iz++, f++, o++) {
                    const R ps = LD(PS[o]), pr = LD(PR[o]);
                    a[f] = a[f] + kr * ps;
                    b[f] = b[f] + ki * ps;
                    c[f] = c[f] + kr * pr;
                    d[f] = d[f] + ki * pr;
                    img[o] = img[o] + w * (a[f] * c[f] - b[f] * d[f]);
                }
            }
    }
}

static W *zalloc(size_t n)
{
    return (W *)calloc(n, sizeof(W));
}

/* rtm-3d/src/kernel/RTMFiniteDifferencesKernel.cpp:129-246 (modeling loop):
 * swap, taperAll(P), taperAll(PP), step, source, save receivers.  rcv is the
 * concatenation of the ntstep chunks, [nt/ntstep][nrx][nry][ntstep].
 * (sx,sy,sz) are interior coordinates. */
int FN(orc_model_shot)(const R *v2, const R *coefs, const R *g, const R *src, int nxe, int nye,
                       int nze, int blen, int order, int nt, int ntstep, int sx, int sy, int sz,
                       const orc_rcv_geom *rg, R *rcv)
{
    const size_t E = (size_t)nxe * nye * nze;
    const size_t chunk = (size_t)rg->count_x * rg->count_y * ntstep;
    W *A = zalloc(E), *B = zalloc(E);
    W *P = A, *PP = B;
    int rc = 0;
    if (!A || !B) { free(A); free(B); return -2; }
    for (int it = 0; it < nt && rc == 0; it++) {
        W *t = P; P = PP; PP = t;
        FN(orc_taper_all)(P, g, nxe, nye, nze, blen);
        FN(orc_taper_all)(PP, g, nxe, nye, nze, blen);
        rc = FN(orc_step)(P, PP, v2, coefs, nxe, nye, nze, order);
        FN(orc_apply_source)(PP, nxe, nye, nze, sx + blen, sy + blen, sz + blen, src[it]);
        FN(orc_receivers)(PP, rcv + (size_t)(it / ntstep) * chunk, nxe, nye, nze, blen, rg,
                          ntstep, it % ntstep, 0);
    }
    free(A);
    free(B);
    return rc;
}

/* HBC migration, rtm-3d/src/kernel/RTMFiniteDifferencesKernel_HBC.cpp:96-365,
 * or RBC migration (hbc == 0), rtm-3d/src/kernel/RTMFiniteDifferencesKernel_RBC.cpp:54-274.
 * Forward: swap, [taperUpper x2], step, source, [save UPB].  Backward, it
 * from nt-1 down to 0: seed from the last two forward fields at nt-1/nt-2,
 * otherwise step the source field back and [restore UPB]; swap; [taperUpper
 * receiver x2]; step receiver; restore receivers (set); image; swap.
 * img is the extended-grid shot image (zeroed here). */
int FN(orc_migrate_shot)(const R *v2, const R *coefs, const R *g, const R *src, int nxe,
                         int nye, int nze, int blen, int order, int nt, int ntstep, int sx,
                         int sy, int sz, const orc_rcv_geom *rg, const R *rcv, int hbc, R *img,
                         W *fwd_last)
{
    const size_t E = (size_t)nxe * nye * nze;
    const size_t chunk = (size_t)rg->count_x * rg->count_y * ntstep;
    const int hf = order / 2;
    W *A = zalloc(E), *B = zalloc(E), *C = zalloc(E), *D = zalloc(E);
    W *upb = hbc ? zalloc((size_t)nt * nxe * nye * hf) : NULL;
    W *snap0, *snap1;
    int rc = 0;
    if (!A || !B || !C || !D || (hbc && !upb)) rc = -2;
    W *P = A, *PP = B;
    for (int it = 0; it < nt && rc == 0; it++) {
        W *t = P; P = PP; PP = t;
        if (hbc) {
            FN(orc_taper_upper)(P, g, nxe, nye, nze, blen);
            FN(orc_taper_upper)(PP, g, nxe, nye, nze, blen);
        }
        rc = FN(orc_step)(P, PP, v2, coefs, nxe, nye, nze, order);
        FN(orc_apply_source)(PP, nxe, nye, nze, sx + blen, sy + blen, sz + blen, src[it]);
        if (hbc) FN(orc_upb)(PP, upb, nxe, nye, nze, blen, order, it, 0);
    }
    snap0 = P;  /* p(nt-2) after its forward taper */
    snap1 = PP; /* p(nt-1) */
    if (fwd_last && rc == 0) memcpy(fwd_last, snap1, E * sizeof(W));
    /* Backward: C/D hold the receiver field; the source field reuses A/B via
     * the snapshot seeding (no copies needed). */
    memset(img, 0, E * sizeof(R));
    W *Ps = snap1, *PPs = snap0;
    W *Pr = C, *PPr = D;
    for (int it = nt - 1; it >= 0 && rc == 0; it--) {
        if (it == nt - 1) {
            /* ppSrc <- snap1, swap: Ps holds p(nt-1); PPs holds zeros that
             * are overwritten at nt-2 before any use. */
            Ps = snap1;
        } else if (it == nt - 2) {
            /* ppSrc <- snap0, swap: Ps = p(nt-2), PPs = p(nt-1). */
            Ps = snap0;
            PPs = snap1;
        } else {
            rc = FN(orc_step)(Ps, PPs, v2, coefs, nxe, nye, nze, order);
            if (hbc) FN(orc_upb)(PPs, upb, nxe, nye, nze, blen, order, it, 1);
            { W *t = Ps; Ps = PPs; PPs = t; }
        }
        if (hbc) {
            FN(orc_taper_upper)(Pr, g, nxe, nye, nze, blen);
            FN(orc_taper_upper)(PPr, g, nxe, nye, nze, blen);
        }
        if (rc == 0) rc = FN(orc_step)(Pr, PPr, v2, coefs, nxe, nye, nze, order);
        FN(orc_receivers)(PPr, (R *)rcv + (size_t)(it / ntstep) * chunk, nxe, nye, nze, blen, rg,
                          ntstep, it % ntstep, 1);
        FN(orc_image)(img, Ps, PPr, nxe, nye, nze, blen);
        { W *t = Pr; Pr = PPr; PPr = t; }
    }
    free(A); free(B); free(C); free(D); free(upb);
    return rc;
}

/* RTMFreqDomainKernel::rtmFreqDomainPropagation
 * (rtm-3d/src/kernel/RTMFiniteDifferencesKernel_FreqDomain.cpp:220-357,
 * :358-369): for every group of nwstep frequencies, from zero fields: per
 * receiver chunk (last first) and per step lt descending with rt = itstep +
 * lt, st = nt-1-rt: step of both fields (rtmStepMultipleWave), taper all
 * borders of the four grids, source at st into PPs, receivers restored
 * (set) into PPr at lt, frequency contributions and image at st, swaps.
 * The image accumulates over the groups. */
int FN(orc_fmig_shot)(const R *v2, const R *coefs, const R *g, const R *src, int nxe, int nye,
                      int nze, int blen, int order, int nt, int ntstep, int sx, int sy, int sz,
                      const orc_rcv_geom *rg, const R *rcv, const R *kre, const R *kim,
                      const R *w2, int nw, int nwstep, R *img)
{
    const size_t E = (size_t)nxe * nye * nze;
    const size_t vol = (size_t)(nxe - 2 * blen) * (nye - 2 * blen) * (nze - 2 * blen);
    const size_t chunk = (size_t)rg->count_x * rg->count_y * ntstep;
    if (nwstep < 1) return -2;
    W *f[4];
    R *acc[4];
    for (int i = 0; i < 4; i++) {
        f[i] = zalloc(E);
        acc[i] = (R *)calloc(vol * (size_t)nwstep, sizeof(R));
    }
    int rc = 0;
    for (int gw = 0; gw < nw && rc == 0; gw += nwstep) {
        for (int i = 0; i < 4; i++) {
            memset(f[i], 0, E * sizeof(W));
            memset(acc[i], 0, vol * (size_t)nwstep * sizeof(R));
        }
        W *Ps = f[0], *PPs = f[1], *Pr = f[2], *PPr = f[3];
        for (int itstep = nt - ntstep; itstep >= 0 && rc == 0; itstep -= ntstep) {
            for (int lt = ntstep - 1; lt >= 0; lt--) {
                const int rt = itstep + lt, st = nt - 1 - rt;
                rc = FN(orc_step)(Ps, PPs, v2, coefs, nxe, nye, nze, order);
                if (FN(orc_step)(Pr, PPr, v2, coefs, nxe, nye, nze, order)) rc = -1;
                if (rc) break;
                FN(orc_taper_all)(Ps, g, nxe, nye, nze, blen);
                FN(orc_taper_all)(PPs, g, nxe, nye, nze, blen);
                FN(orc_taper_all)(Pr, g, nxe, nye, nze, blen);
                FN(orc_taper_all)(PPr, g, nxe, nye, nze, blen);
                FN(orc_apply_source)(PPs, nxe, nye, nze, sx + blen, sy + blen, sz + blen, src[st]);
                FN(orc_receivers)(PPr, (R *)rcv + (size_t)(itstep / ntstep) * chunk, nxe, nye,
                                  nze, blen, rg, ntstep, lt, 1);
                FN(orc_freq_image)(img, PPs, PPr, acc[0], acc[1], acc[2], acc[3], kre, kim, w2,
                                   nxe, nye, nze, blen, nwstep, gw, nw, nt, st);
                W *t = Ps;
                Ps = PPs;
                PPs = t;
                t = Pr;
                Pr = PPr;
                PPr = t;
            }
        }
    }
    for (int i = 0; i < 4; i++) {
        free(f[i]);
        free(acc[i]);
    }
    return rc;
}
