// rtm_stencil.hip -- second-order-in-time star-stencil update for gfx950.
//
// Restates rtm-3d/src/platform/host/RTMCPUPlatform.cpp:325-395 (rtmStep3D):
// for x,y,z in [R, n-R), acm_d = sum_{io=0..2R} P[..i_d-R+io..]*c_d[io]
// accumulated from 0 in ascending io, lapl = (acmx+acmy)+acmz,
// PP <- ((2p) - pp) + v2dt2*lapl in place.  Compiled with -ffp-contract=off
// (and the pragma below) so every product and sum is rounded separately, as
// in the reference, which is built without FMA.
//
// Mapping (grid layout [x][y][z], z fastest):
//   * a block owns a TZ-wide z range (z on lanes: coalesced 256 B rows), a
//     TY = BY*RY row range in y, and marches a chunk of x planes;
//   * each thread keeps its RY cells' x-neighbourhood in a (2R+1)-deep
//     register queue that rotates by one plane per step;
//   * the centre plane with its y and z halos is staged in LDS (double
//     buffered: one barrier per plane) for the y and z taps;
//   * the next plane's loads (queue lead, halos, PP, V) are issued before the
//     current plane's arithmetic so memory latency overlaps compute;
//   * block ids are remapped so that consecutive tiles (which share halo
//     rows) land on the same XCD and therefore the same L2.
// Optional epilogue (IMG): the imaging condition img += ps*pp of the
// backward receiver step (RTMCPUPlatform.cpp:513-540) for z in [z0, z1),
// skipping receiver cells (those are imaged after receiver injection).
#include "rtm_internal.hpp"

#include <algorithm>

#pragma clang fp contract(off)

namespace rtmhip {
namespace {

struct TileGrid {
    int ntz, nty, nxc, xchunk;
};

// Hardware dispatches workgroup h to XCD h % 8; give XCD k a contiguous range
// of logical tiles (z fastest, then y, then x chunk).
__device__ __forceinline__ void block_coords(const TileGrid &tg, int &bz, int &by, int &bx)
{
    const int n = tg.ntz * tg.nty * tg.nxc;
    const int h = blockIdx.x;
    const int q = n >> 3, r = n & 7, k = h & 7, j = h >> 3;
    const int t = k * q + min(k, r) + j;
    bz = t % tg.ntz;
    const int t2 = t / tg.ntz;
    by = t2 % tg.nty;
    bx = t2 / tg.nty;
}

template <int R, int TZ, int BY, int RY, bool IMG>
__global__ void __launch_bounds__(TZ *BY)
    k_stencil(const float *__restrict__ P, float *__restrict__ PP, const float *__restrict__ V,
              const Coefs cf, const Grid g, const TileGrid tg, const ImageFuse fz,
              int *__restrict__ flag)
{
    constexpr int TY = BY * RY;
    constexpr int LW = TZ + 2 * R;
    constexpr int LH = TY + 2 * R;
    constexpr int NH = 2 * R;
    constexpr int NYH = (NH + BY - 1) / BY;
    constexpr int NZH = (TY * NH + TZ * BY - 1) / (TZ * BY);
    __shared__ float tile[2][LH][LW];

    int bz, by, bx;
    block_coords(tg, bz, by, bx);
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * TZ + tx;
    const int z0 = bz * TZ, y0 = R + by * TY;
    const int z = z0 + tx;
    const int xb = R + bx * tg.xchunk;
    const int xe = min(xb + tg.xchunk, g.nx - R);
    const bool zload = z < g.nz;
    const bool zcomp = (z >= R) && (z < g.nz - R);

    // Own rows, interleaved across thread rows.
    int roff[RY];
    bool yload[RY], ycomp[RY], rcvrow[RY];
#pragma unroll
    for (int r = 0; r < RY; ++r) {
        const int y = y0 + ty + BY * r;
        roff[r] = y * g.sy + z;
        yload[r] = zload && (y < g.ny);
        ycomp[r] = zcomp && (y < g.ny - R);
        rcvrow[r] = false;
        if (IMG) {
            const int dy = y - fz.ry0;
            rcvrow[r] = (z == fz.rz) && dy >= 0 && (dy % fz.rdy) == 0 && (dy / fz.rdy) < fz.rny;
        }
    }
    bool imgz = false;
    if (IMG) imgz = (z >= fz.z0) && (z < fz.z1);

    // y halo: R rows above and R rows below the tile, all TZ lanes.
    int yh_off[NYH], yh_lrow[NYH];
    bool yh_ok[NYH], yh_valid[NYH];
#pragma unroll
    for (int i = 0; i < NYH; ++i) {
        const int h = ty + BY * i;
        yh_valid[i] = h < NH;
        const int gy = h < R ? y0 - R + h : y0 + TY + (h - R);
        yh_lrow[i] = h < R ? h : R + TY + (h - R);
        yh_ok[i] = yh_valid[i] && zload && gy < g.ny;
        yh_off[i] = gy * g.sy + z;
    }
    // z halo: R columns left and right of the tile, all TY rows.
    int zh_off[NZH], zh_lrow[NZH], zh_lcol[NZH];
    bool zh_ok[NZH], zh_valid[NZH];
#pragma unroll
    for (int k = 0; k < NZH; ++k) {
        const int i = tid + TZ * BY * k;
        zh_valid[k] = i < TY * NH;
        const int rrow = i / NH, c = i % NH;
        const int zz = c < R ? z0 - R + c : z0 + TZ + (c - R);
        zh_lcol[k] = c < R ? c : R + TZ + (c - R);
        zh_lrow[k] = R + rrow;
        const int gy = y0 + rrow;
        zh_ok[k] = zh_valid[k] && zz >= 0 && zz < g.nz && gy < g.ny;
        zh_off[k] = gy * g.sy + zz;
    }

    float q[2 * R + 1][RY];
#pragma unroll
    for (int k = 0; k < 2 * R; ++k) {
        const long long pl = (long long)(xb - R + k) * g.sx;
#pragma unroll
        for (int r = 0; r < RY; ++r) q[k][r] = yload[r] ? P[pl + roff[r]] : 0.f;
    }

    float n_lead[RY], n_pp[RY], n_v[RY], n_ps[RY], n_img[RY], n_yh[NYH], n_zh[NZH];
    auto fetch = [&](int x) {
        const long long pl = (long long)x * g.sx;
        const long long ld = (long long)(x + R) * g.sx;
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            n_lead[r] = yload[r] ? P[ld + roff[r]] : 0.f;
            n_pp[r] = ycomp[r] ? PP[pl + roff[r]] : 0.f;
            n_v[r] = ycomp[r] ? V[pl + roff[r]] : 0.f;
            if (IMG) {
                const bool im = ycomp[r] && imgz;
                n_ps[r] = im ? fz.ps[pl + roff[r]] : 0.f;
                n_img[r] = im ? fz.img[pl + roff[r]] : 0.f;
            }
        }
#pragma unroll
        for (int i = 0; i < NYH; ++i) n_yh[i] = yh_ok[i] ? P[pl + yh_off[i]] : 0.f;
#pragma unroll
        for (int k = 0; k < NZH; ++k) n_zh[k] = zh_ok[k] ? P[pl + zh_off[k]] : 0.f;
    };

    bool bad = false;
    if (xb < xe) fetch(xb);
    for (int x = xb; x < xe; ++x) {
        const int buf = (x - xb) & 1;
        float c_pp[RY], c_v[RY], c_ps[RY], c_img[RY];
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            q[2 * R][r] = n_lead[r];
            c_pp[r] = n_pp[r];
            c_v[r] = n_v[r];
            if (IMG) {
                c_ps[r] = n_ps[r];
                c_img[r] = n_img[r];
            }
            tile[buf][R + ty + BY * r][R + tx] = q[R][r];
        }
#pragma unroll
        for (int i = 0; i < NYH; ++i)
            if (yh_valid[i]) tile[buf][yh_lrow[i]][R + tx] = n_yh[i];
#pragma unroll
        for (int k = 0; k < NZH; ++k)
            if (zh_valid[k]) tile[buf][zh_lrow[k]][zh_lcol[k]] = n_zh[k];
        if (x + 1 < xe) fetch(x + 1);
        __syncthreads();

        bool rcvx = false;
        if (IMG) {
            const int dx = x - fz.rx0;
            rcvx = dx >= 0 && (dx % fz.rdx) == 0 && (dx / fz.rdx) < fz.rnx;
        }
        const long long pl = (long long)x * g.sx;
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            if (!ycomp[r]) continue;
            const float *t = &tile[buf][R + ty + BY * r][R + tx];
            const float p = q[R][r];
            float ax = 0.f, ay = 0.f, az = 0.f;
#pragma unroll
            for (int io = 0; io <= 2 * R; ++io) {
                const float vy = (io == R) ? p : t[(io - R) * LW];
                const float vz = (io == R) ? p : t[io - R];
                ax += q[io][r] * cf.c[0][io];
                ay += vy * cf.c[1][io];
                az += vz * cf.c[2][io];
            }
            const float lapl = ax + ay + az;
            const float npp = 2.f * p - c_pp[r] + c_v[r] * lapl;
            PP[pl + roff[r]] = npp;
            bad |= (npp != npp);
            if (IMG && imgz && !(rcvx && rcvrow[r]))
                fz.img[pl + roff[r]] = c_img[r] + c_ps[r] * npp;
        }
#pragma unroll
        for (int k = 0; k < 2 * R; ++k)
#pragma unroll
            for (int r = 0; r < RY; ++r) q[k][r] = q[k + 1][r];
    }
    if (bad) atomicOr(flag, 1);
}

template <int R, int TZ, int BY, int RY>
hipError_t launch_cfg(const float *P, float *PP, const float *V, const Coefs &cf, const Grid &g,
                      const ImageFuse *fuse, int *flag, hipStream_t s, int xchunk)
{
    constexpr int TY = BY * RY;
    TileGrid tg;
    tg.ntz = (g.nz + TZ - 1) / TZ;
    tg.nty = (g.ny - 2 * R + TY - 1) / TY;
    const int nxi = g.nx - 2 * R;
    if (tg.nty <= 0 || nxi <= 0 || g.nz - 2 * R <= 0) return hipSuccess;  // nothing to update
    if (xchunk <= 0) {
        // Aim for about 2048 blocks (8 per CU) so every XCD stays busy, but
        // march at least 16 planes per block to amortise the queue fill.
        const long long tiles = (long long)tg.ntz * tg.nty;
        long long want = (2048 + tiles - 1) / tiles;
        if (want < 1) want = 1;
        xchunk = (int)((nxi + want - 1) / want);
        if (xchunk < 16) xchunk = 16;
    }
    tg.xchunk = xchunk;
    tg.nxc = (nxi + xchunk - 1) / xchunk;
    const dim3 block(TZ, BY);
    const dim3 grid((unsigned)(tg.ntz * tg.nty * tg.nxc));
    ImageFuse fz{};
    if (fuse) {
        fz = *fuse;
        if (fz.rdx <= 0) fz.rdx = 1;
        if (fz.rdy <= 0) fz.rdy = 1;
        hipLaunchKernelGGL((k_stencil<R, TZ, BY, RY, true>), grid, block, 0, s, P, PP, V, cf, g,
                           tg, fz, flag);
    } else {
        hipLaunchKernelGGL((k_stencil<R, TZ, BY, RY, false>), grid, block, 0, s, P, PP, V, cf,
                           g, tg, fz, flag);
    }
    return hipGetLastError();
}

template <int R>
hipError_t launch_radius(const float *P, float *PP, const float *V, const Coefs &cf,
                         const Grid &g, const ImageFuse *fuse, int *flag, hipStream_t s,
                         const StencilTiling *t)
{
    const int xchunk = t ? t->xchunk : 0;
    if (!t || (t->tz == 64 && t->by == 4 && t->ry == 4))
        return launch_cfg<R, 64, 4, 4>(P, PP, V, cf, g, fuse, flag, s, xchunk);
    if (t->tz == 64 && t->by == 4 && t->ry == 2)
        return launch_cfg<R, 64, 4, 2>(P, PP, V, cf, g, fuse, flag, s, xchunk);
    if (t->tz == 64 && t->by == 4 && t->ry == 8)
        return launch_cfg<R, 64, 4, 8>(P, PP, V, cf, g, fuse, flag, s, xchunk);
    if (t->tz == 64 && t->by == 8 && t->ry == 4)
        return launch_cfg<R, 64, 8, 4>(P, PP, V, cf, g, fuse, flag, s, xchunk);
    if (t->tz == 32 && t->by == 8 && t->ry == 4)
        return launch_cfg<R, 32, 8, 4>(P, PP, V, cf, g, fuse, flag, s, xchunk);
    return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_stencil(int radius, const float *P, float *PP, const float *V, const Coefs &cf,
                          const Grid &g, const ImageFuse *fuse, int *flag, hipStream_t s,
                          const StencilTiling *tiling)
{
    switch (radius) {
    case 1: return launch_radius<1>(P, PP, V, cf, g, fuse, flag, s, tiling);
    case 2: return launch_radius<2>(P, PP, V, cf, g, fuse, flag, s, tiling);
    case 3: return launch_radius<3>(P, PP, V, cf, g, fuse, flag, s, tiling);
    case 4: return launch_radius<4>(P, PP, V, cf, g, fuse, flag, s, tiling);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace rtmhip
