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
//   * a block owns a TZ-wide z range (z on lanes: coalesced rows), a
//     TY = BY*RY row range in y, and marches a chunk of x planes;
//   * each thread keeps its RY cells' x-neighbourhood in a ring of 2R+2
//     register planes; the x loop is unrolled by the ring length so that
//     ring slots are compile-time registers (no data movement per plane);
//   * the centre plane with its y and z halos is staged in LDS (double
//     buffered: one barrier per plane) for the y and z taps;
//   * the next plane's loads (ring lead, halos, PP, V, imaging operands) are
//     issued before the current plane's arithmetic (ping-pong registers);
//   * out-of-grid lanes load a clamped in-grid address and never use the
//     value, so loads carry no per-lane branches; offsets inside a plane are
//     32-bit with the plane base in scalar registers;
//   * block ids are remapped so that consecutive tiles (which share halo
//     rows) land on the same XCD and therefore the same L2.
// Optional fusions:
//   IMG   -- the imaging condition img += ps*pp of the backward receiver step
//            (RTMCPUPlatform.cpp:513-540) for z in [z0, z1), skipping
//            receiver cells (imaged after receiver injection);
//   TAPER -- the upper-border taper of both wavefields that the reference
//            applies in place before each HBC step (taperUpper3D,
//            RTMCPUPlatform.cpp:566-657): P values are tapered once as they
//            are loaded and the centre PP value pp_taps times, with the
//            reference's factor order; memory is left untapered and the
//            caller applies the one pending in-place taper at the end.
#include "rtm_internal.hpp"

#include <algorithm>
#include <map>
#include <mutex>
#include <type_traits>
#include <utility>

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

// One application of the upper taper to a value at (x, y, z): nothing below
// z = b; above it ((v*fy)*fx)*fz with fy, fx = 1 off the lateral borders
// (multiplying by 1 is exact, so this is the reference's per-region order).
__device__ __forceinline__ float taper1(float v, float fx, float fy, float fz)
{
    v *= fy;
    v *= fx;
    v *= fz;
    return v;
}

__device__ __forceinline__ float lat_factor(const float *__restrict__ t, int i, int n, int b)
{
    return i < b ? t[b - 1 - i] : (i >= n - b ? t[i - (n - b)] : 1.0f);
}

template <int R, int TZ, int BY, int RY, bool IMG, bool TAPER>
__global__ void __launch_bounds__(TZ *BY)
    k_stencil(const float *__restrict__ P, float *__restrict__ PP, const float *__restrict__ V,
              const Coefs cf, const Grid g, const TileGrid tg, const ImageFuse fz,
              const TaperFuse tf, int *__restrict__ flag)
{
    constexpr int TY = BY * RY;
    constexpr int LW = TZ + 2 * R;
    constexpr int LH = TY + 2 * R;
    constexpr int NH = 2 * R;
    constexpr int NYH = (NH + BY - 1) / BY;
    constexpr int NZH = (TY * NH + TZ * BY - 1) / (TZ * BY);
    constexpr int NR = 2 * R + 2;  // ring length (even)
    __shared__ float tile[2][LH][LW];

    int bz, by, bx;
    block_coords(tg, bz, by, bx);
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * TZ + tx;
    const int z0 = bz * TZ, y0 = R + by * TY;
    const int z = z0 + tx;
    const int xb = R + bx * tg.xchunk;
    const int xe = min(xb + tg.xchunk, g.nx - R);
    const int zc = min(z, g.nz - 1);
    const bool zcomp = (z >= R) && (z < g.nz - R);
    const unsigned sy = (unsigned)g.sy;

    // Taper applies to this block only if it loads any z < b.
    const bool btap = TAPER && (z0 - R < tf.b);
    float fzl = 1.0f;  // own lane's z factor
    if (TAPER && btap && z < tf.b) fzl = tf.t[tf.b - 1 - z];

    unsigned roff[RY];
    bool ycomp[RY], rcvrow[RY];
    float fyr[RY];
#pragma unroll
    for (int r = 0; r < RY; ++r) {
        const int y = y0 + ty + BY * r;
        roff[r] = (unsigned)min(y, g.ny - 1) * sy + (unsigned)zc;
        ycomp[r] = zcomp && (y < g.ny - R);
        rcvrow[r] = false;
        if (IMG) {
            const int dy = y - fz.ry0;
            rcvrow[r] = (z == fz.rz) && dy >= 0 && (dy % fz.rdy) == 0 && (dy / fz.rdy) < fz.rny;
        }
        fyr[r] = 1.0f;
        if (TAPER && btap && z < tf.b) fyr[r] = lat_factor(tf.t, min(y, g.ny - 1), g.ny, tf.b);
    }
    bool imgz = false;
    if (IMG) imgz = (z >= fz.z0) && (z < fz.z1);

    // y halo: R rows above and R rows below the tile, all TZ lanes.
    unsigned yh_off[NYH];
    int yh_lrow[NYH];
    bool yh_valid[NYH];
    float yh_fy[NYH];
#pragma unroll
    for (int i = 0; i < NYH; ++i) {
        const int h = ty + BY * i;
        yh_valid[i] = h < NH;
        const int gy = min(h < R ? y0 - R + h : y0 + TY + (h - R), g.ny - 1);
        yh_lrow[i] = h < R ? h : R + TY + (h - R);
        yh_off[i] = (unsigned)gy * sy + (unsigned)zc;
        yh_fy[i] = 1.0f;
        if (TAPER && btap && z < tf.b) yh_fy[i] = lat_factor(tf.t, gy, g.ny, tf.b);
    }
    // z halo: R columns left and right of the tile, all TY rows.
    unsigned zh_off[NZH];
    int zh_lrow[NZH], zh_lcol[NZH];
    bool zh_valid[NZH], zh_tap[NZH];
    float zh_fy[NZH], zh_fz[NZH];
#pragma unroll
    for (int k = 0; k < NZH; ++k) {
        const int i = tid + TZ * BY * k;
        zh_valid[k] = i < TY * NH;
        const int rrow = i / NH, c = i % NH;
        const int zz = min(max(c < R ? z0 - R + c : z0 + TZ + (c - R), 0), g.nz - 1);
        zh_lcol[k] = c < R ? c : R + TZ + (c - R);
        zh_lrow[k] = R + rrow;
        const int gy = min(y0 + rrow, g.ny - 1);
        zh_off[k] = (unsigned)gy * sy + (unsigned)zz;
        zh_fy[k] = 1.0f;
        zh_fz[k] = 1.0f;
        zh_tap[k] = TAPER && btap && zz < tf.b;
        if (zh_tap[k]) {
            zh_fy[k] = lat_factor(tf.t, gy, g.ny, tf.b);
            zh_fz[k] = tf.t[tf.b - 1 - zz];
        }
    }

    // Ring of planes: slot s holds plane xb - R + s (mod NR).
    float q[NR][RY];
    float n_pp[2][RY], n_v[2][RY], n_ps[2][RY], n_img[2][RY], n_yh[2][NYH], n_zh[2][NZH];

    auto plane_fx = [&](int x) -> float {
        return (TAPER && btap) ? lat_factor(tf.t, x, g.nx, tf.b) : 1.0f;
    };
    // Loads of P at plane x for the ring slot `slot` (own rows).
    auto load_ring = [&](int x, float(&dst)[RY]) {
        const float *__restrict__ Px = P + (size_t)x * (size_t)g.sx;
        const float fx = plane_fx(x);
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            float v = Px[roff[r]];
            if (TAPER && btap) v = taper1(v, z < tf.b ? fx : 1.0f, fyr[r], fzl);
            dst[r] = v;
        }
    };
    auto fetch = [&](int x, int set) {
        const size_t pl = (size_t)x * (size_t)g.sx;
        const float *__restrict__ Px = P + pl;
        const float fx = plane_fx(x);
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            n_pp[set][r] = PP[pl + roff[r]];
            n_v[set][r] = V[pl + roff[r]];
            if (IMG) {
                n_ps[set][r] = fz.ps[pl + roff[r]];
                n_img[set][r] = fz.img[pl + roff[r]];
            }
        }
#pragma unroll
        for (int i = 0; i < NYH; ++i) {
            float v = Px[yh_off[i]];
            if (TAPER && btap) v = taper1(v, z < tf.b ? fx : 1.0f, yh_fy[i], fzl);
            n_yh[set][i] = v;
        }
#pragma unroll
        for (int k = 0; k < NZH; ++k) {
            float v = Px[zh_off[k]];
            if (TAPER && btap) v = taper1(v, zh_tap[k] ? fx : 1.0f, zh_fy[k], zh_fz[k]);
            n_zh[set][k] = v;
        }
    };

#pragma unroll
    for (int s = 0; s < 2 * R; ++s) load_ring(xb - R + s, q[s]);
    load_ring(xb + R, q[2 * R]);
    fetch(xb, 0);

    bool bad = false;
    int x = xb;
    // One plane; K = (x - xb) mod NR is a compile-time constant so every ring
    // slot and ping-pong set below is a fixed register.  Returns false once
    // the march is done.
    auto step = [&]<int K>(std::integral_constant<int, K>) -> bool {
        if (x >= xe) return false;
        constexpr int cur = K & 1;
        constexpr int ctr = (K + R) % NR;
        // stage the centre plane and its halos
#pragma unroll
        for (int r = 0; r < RY; ++r) tile[cur][R + ty + BY * r][R + tx] = q[ctr][r];
#pragma unroll
        for (int i = 0; i < NYH; ++i)
            if (yh_valid[i]) tile[cur][yh_lrow[i]][R + tx] = n_yh[cur][i];
#pragma unroll
        for (int j = 0; j < NZH; ++j)
            if (zh_valid[j]) tile[cur][zh_lrow[j]][zh_lcol[j]] = n_zh[cur][j];
        // next plane's loads
        if (x + 1 < xe) {
            load_ring(x + 1 + R, q[(K + 2 * R + 1) % NR]);
            fetch(x + 1, cur ^ 1);
        }
        __syncthreads();

        bool rcvx = false;
        if (IMG) {
            const int dx = x - fz.rx0;
            rcvx = dx >= 0 && (dx % fz.rdx) == 0 && (dx / fz.rdx) < fz.rnx;
        }
        const size_t pl = (size_t)x * (size_t)g.sx;
        float fxc = 1.0f;
        if (TAPER && btap) fxc = plane_fx(x);
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            const float *t = &tile[cur][R + ty + BY * r][R + tx];
            const float p = q[ctr][r];
            float ax = 0.f, ay = 0.f, az = 0.f;
#pragma unroll
            for (int io = 0; io <= 2 * R; ++io) {
                const float vy = (io == R) ? p : t[(io - R) * LW];
                const float vz = (io == R) ? p : t[io - R];
                ax += q[(K + io) % NR][r] * cf.c[0][io];
                ay += vy * cf.c[1][io];
                az += vz * cf.c[2][io];
            }
            const float lapl = ax + ay + az;
            float pp = n_pp[cur][r];
            if (TAPER && btap && z < tf.b) {
                pp = taper1(pp, fxc, fyr[r], fzl);
                if (tf.pp_taps > 1) pp = taper1(pp, fxc, fyr[r], fzl);
            }
            const float npp = 2.f * p - pp + n_v[cur][r] * lapl;
            if (ycomp[r]) {
                PP[pl + roff[r]] = npp;
                bad |= (npp != npp);
                if (IMG && imgz && !(rcvx && rcvrow[r]))
                    fz.img[pl + roff[r]] = n_img[cur][r] + n_ps[cur][r] * npp;
            }
        }
        ++x;
        return true;
    };
    while (x < xe) {
        [&]<int... Ks>(std::integer_sequence<int, Ks...>) {
            (step(std::integral_constant<int, Ks>{}) && ...);
        }(std::make_integer_sequence<int, NR>{});
    }
    if (bad) atomicOr(flag, 1);
}

struct LaunchKey {
    const void *fn;
    int threads;
    bool operator<(const LaunchKey &o) const
    {
        return fn < o.fn || (fn == o.fn && threads < o.threads);
    }
};

// Resident blocks on the whole device for one kernel (occupancy x CUs).
int device_capacity(const void *fn, int threads)
{
    static std::mutex mu;
    static std::map<LaunchKey, int> cache;
    std::lock_guard<std::mutex> lock(mu);
    const LaunchKey key{fn, threads};
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int dev = 0, ncu = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        ncu = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, 0) != hipSuccess ||
        per_cu <= 0)
        per_cu = 4;
    const int cap = std::max(1, ncu * per_cu);
    cache[key] = cap;
    return cap;
}

// Pick the number of x chunks per tile column: whole rounds of resident
// blocks (no half-empty last round) and long marches (the ring preload of
// 2R planes per chunk is overhead).
int choose_chunks(long long tiles, int nxi, int R, int cap)
{
    double best = -1.0;
    int best_n = 1;
    for (int n = 1; n <= 64 && n <= nxi; ++n) {
        const long long total = tiles * n;
        const long long rounds = (total + cap - 1) / cap;
        const int xc = (nxi + n - 1) / n;
        const double fill = (double)total / (double)(rounds * cap);
        const double march = (double)xc / (double)(xc + 2 * R);
        const double eff = fill * march;
        if (eff > best + 1e-9) {
            best = eff;
            best_n = n;
        }
    }
    return best_n;
}

template <int R, int TZ, int BY, int RY, bool IMG, bool TAPER>
hipError_t launch_cfg(const float *P, float *PP, const float *V, const Coefs &cf, const Grid &g,
                      const ImageFuse &fz, const TaperFuse &tf, int *flag, hipStream_t s,
                      int xchunk)
{
    constexpr int TY = BY * RY;
    auto kern = k_stencil<R, TZ, BY, RY, IMG, TAPER>;
    TileGrid tg;
    tg.ntz = (g.nz + TZ - 1) / TZ;
    tg.nty = (g.ny - 2 * R + TY - 1) / TY;
    const int nxi = g.nx - 2 * R;
    if (tg.nty <= 0 || nxi <= 0 || g.nz - 2 * R <= 0) return hipSuccess;  // nothing to update
    if (xchunk <= 0) {
        const int cap = device_capacity((const void *)kern, TZ * BY);
        const int n = choose_chunks((long long)tg.ntz * tg.nty, nxi, R, cap);
        xchunk = (nxi + n - 1) / n;
    }
    tg.xchunk = xchunk;
    tg.nxc = (nxi + xchunk - 1) / xchunk;
    const dim3 block(TZ, BY);
    const dim3 grid((unsigned)(tg.ntz * tg.nty * tg.nxc));
    hipLaunchKernelGGL(kern, grid, block, 0, s, P, PP, V, cf, g, tg, fz, tf, flag);
    return hipGetLastError();
}

template <int R, int TZ, int BY, int RY>
hipError_t launch_variant(const float *P, float *PP, const float *V, const Coefs &cf,
                          const Grid &g, const ImageFuse *fuse, const TaperFuse *taper,
                          int *flag, hipStream_t s, int xchunk)
{
    ImageFuse fz{};
    TaperFuse tf{};
    if (fuse) {
        fz = *fuse;
        if (fz.rdx <= 0) fz.rdx = 1;
        if (fz.rdy <= 0) fz.rdy = 1;
    }
    if (taper) tf = *taper;
    const bool tap = taper && taper->b > 0;
    if (fuse && tap)
        return launch_cfg<R, TZ, BY, RY, true, true>(P, PP, V, cf, g, fz, tf, flag, s, xchunk);
    if (fuse)
        return launch_cfg<R, TZ, BY, RY, true, false>(P, PP, V, cf, g, fz, tf, flag, s, xchunk);
    if (tap)
        return launch_cfg<R, TZ, BY, RY, false, true>(P, PP, V, cf, g, fz, tf, flag, s, xchunk);
    return launch_cfg<R, TZ, BY, RY, false, false>(P, PP, V, cf, g, fz, tf, flag, s, xchunk);
}

template <int R>
hipError_t launch_radius(const float *P, float *PP, const float *V, const Coefs &cf,
                         const Grid &g, const ImageFuse *fuse, const TaperFuse *taper, int *flag,
                         hipStream_t s, const StencilTiling *t)
{
    const int xchunk = t ? t->xchunk : 0;
    if (!t || (t->tz == 64 && t->by == 4 && t->ry == 4))
        return launch_variant<R, 64, 4, 4>(P, PP, V, cf, g, fuse, taper, flag, s, xchunk);
    if (t->tz == 64 && t->by == 4 && t->ry == 2)
        return launch_variant<R, 64, 4, 2>(P, PP, V, cf, g, fuse, taper, flag, s, xchunk);
    if (t->tz == 64 && t->by == 8 && t->ry == 2)
        return launch_variant<R, 64, 8, 2>(P, PP, V, cf, g, fuse, taper, flag, s, xchunk);
    if (t->tz == 32 && t->by == 8 && t->ry == 4)
        return launch_variant<R, 32, 8, 4>(P, PP, V, cf, g, fuse, taper, flag, s, xchunk);
    if (t->tz == 32 && t->by == 8 && t->ry == 2)
        return launch_variant<R, 32, 8, 2>(P, PP, V, cf, g, fuse, taper, flag, s, xchunk);
    return hipErrorInvalidValue;
}

}  // namespace

bool stencil_tiling_supported(int tz, int by, int ry)
{
    return (tz == 64 && by == 4 && (ry == 4 || ry == 2)) || (tz == 64 && by == 8 && ry == 2) ||
           (tz == 32 && by == 8 && (ry == 4 || ry == 2));
}

hipError_t launch_stencil(int radius, const float *P, float *PP, const float *V, const Coefs &cf,
                          const Grid &g, const ImageFuse *fuse, const TaperFuse *taper,
                          int *flag, hipStream_t s, const StencilTiling *tiling)
{
    switch (radius) {
    case 1: return launch_radius<1>(P, PP, V, cf, g, fuse, taper, flag, s, tiling);
    case 2: return launch_radius<2>(P, PP, V, cf, g, fuse, taper, flag, s, tiling);
    case 3: return launch_radius<3>(P, PP, V, cf, g, fuse, taper, flag, s, tiling);
    case 4: return launch_radius<4>(P, PP, V, cf, g, fuse, taper, flag, s, tiling);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace rtmhip
