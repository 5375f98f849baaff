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
#include <atomic>
#include <cstdint>
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

// PP, V and the imaging operands are touched once per step; marking them
// non-temporal keeps L2 capacity for P, whose rows are re-read as halos by
// neighbouring tiles a few planes after their owner loaded them.
// (Measured on MI355X: non-temporal hints made the step slower, so they are
// off unless RTM_STREAM_NT=1 is defined at build time.)
#ifndef RTM_STREAM_NT
#define RTM_STREAM_NT 0
#endif
template <typename T>
__device__ __forceinline__ float stream_load(const T *p)
{
#if RTM_STREAM_NT
    return wld(__builtin_nontemporal_load(p));
#else
    return wld(*p);
#endif
}

template <typename T>
__device__ __forceinline__ void stream_store(T *p, float v)
{
#if RTM_STREAM_NT
    __builtin_nontemporal_store(wst<T>(v), p);
#else
    *p = wst<T>(v);
#endif
}

__device__ __forceinline__ float lat_factor(const float *__restrict__ t, int i, int n, int b)
{
    return i < b ? t[b - 1 - i] : (i >= n - b ? t[i - (n - b)] : 1.0f);
}

template <typename T, int R, int TZ, int BY, int RY, bool IMG, bool TAPER>
__global__ void __launch_bounds__(TZ *BY)
    k_stencil(const T *__restrict__ P, T *__restrict__ PP, const float *__restrict__ V,
              const Coefs cf, const Grid g, const TileGrid tg, const ImageFuse fz,
              const TaperFuse tf, int *__restrict__ flag)
{
    const T *__restrict__ PS = static_cast<const T *>(fz.ps);
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
    // A taper applied in registers is one in-place pass of the reference:
    // its result is rounded to the storage type (exact for float).
    auto load_ring = [&](int x, float(&dst)[RY]) {
        const T *__restrict__ Px = P + (size_t)x * (size_t)g.sx;
        const float fx = plane_fx(x);
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            float v = wld(Px[roff[r]]);
            if (TAPER && btap) v = wrnd<T>(taper1(v, z < tf.b ? fx : 1.0f, fyr[r], fzl));
            dst[r] = v;
        }
    };
    auto fetch = [&](int x, int set) {
        const size_t pl = (size_t)x * (size_t)g.sx;
        const T *__restrict__ Px = P + pl;
        const float fx = plane_fx(x);
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            n_pp[set][r] = stream_load(&PP[pl + roff[r]]);
            n_v[set][r] = stream_load(&V[pl + roff[r]]);
            if (IMG) {
                n_ps[set][r] = stream_load(&PS[pl + roff[r]]);
                n_img[set][r] = stream_load(&fz.img[pl + roff[r]]);
            }
        }
#pragma unroll
        for (int i = 0; i < NYH; ++i) {
            float v = wld(Px[yh_off[i]]);
            if (TAPER && btap) v = wrnd<T>(taper1(v, z < tf.b ? fx : 1.0f, yh_fy[i], fzl));
            n_yh[set][i] = v;
        }
#pragma unroll
        for (int k = 0; k < NZH; ++k) {
            float v = wld(Px[zh_off[k]]);
            if (TAPER && btap) v = wrnd<T>(taper1(v, zh_tap[k] ? fx : 1.0f, zh_fy[k], zh_fz[k]));
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
                pp = wrnd<T>(taper1(pp, fxc, fyr[r], fzl));
                if (tf.pp_taps > 1) pp = wrnd<T>(taper1(pp, fxc, fyr[r], fzl));
            }
            const float npp = 2.f * p - pp + n_v[cur][r] * lapl;
            if (ycomp[r]) {
                stream_store(&PP[pl + roff[r]], npp);
                bad |= (npp != npp);
                // the image reads the stored (rounded) receiver value
                if (IMG && imgz && !(rcvx && rcvrow[r]))
                    stream_store(&fz.img[pl + roff[r]],
                                 n_img[cur][r] + n_ps[cur][r] * wrnd<T>(npp));
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

// ---------------------------------------------------------------------------
// Full-row variant.  A block covers every z of its rows: 64*NW lanes per row
// (NW = ceil(nz/64) <= NWM waves), so no z halo is loaded and each row is
// read as one contiguous segment (a 64-wide z tile wastes most of the two
// cache lines its 4-cell z halos touch).  BY thread rows x RY rows per
// thread; the y halo rows are shared through LDS as in the tiled kernel.
// Every wave lies in one thread row, so y and everything derived from it
// (row offsets, halo rows, lateral taper factors, receiver rows) is
// wave-uniform and lives in scalar registers: a global access is a scalar
// row base plus the lane's 32-bit z byte offset, and an LDS access is one
// per-lane base plus a compile-time offset (the LDS row pitch is fixed by
// NWM).  The imaging operands (IMG) are loaded at the start of their own
// plane step rather than a plane ahead, to bound the register count.
struct RowGrid {
    int nty, nxc, xchunk;
};

__device__ __forceinline__ void row_block_coords(const RowGrid &rg, int &by, int &bx)
{
    const int n = rg.nty * rg.nxc;
    const int h = blockIdx.x;
    const int q = n >> 3, r = n & 7, k = h & 7, j = h >> 3;
    const int t = k * q + min(k, r) + j;
    by = t % rg.nty;
    bx = t / rg.nty;
}

// Raw buffer view of one plane: the 128-bit descriptor (base address, no
// stride, range check off, 32-bit data format) lives in scalar registers, so
// an access is descriptor + scalar row byte offset + the lane's 32-bit z byte
// offset -- one VGPR of address per lane for every row and every array.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const void *p)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, -1, 0x00020000);
}
// Loads and stores of one element of type T, widened to / rounded from
// float.  The builtins move raw bits: reinterpret, never convert.
template <typename T>
__device__ __forceinline__ float ld_buf(__amdgpu_buffer_rsrc_t rs, unsigned soff, unsigned voff)
{
    if constexpr (sizeof(T) == 4)
        return __builtin_bit_cast(
            float, __builtin_amdgcn_raw_buffer_load_b32(rs, (int)voff, (int)soff, 0));
    else
        return wld(__builtin_bit_cast(
            T, __builtin_amdgcn_raw_buffer_load_b16(rs, (int)voff, (int)soff, 0)));
}
template <typename T>
__device__ __forceinline__ void st_buf(__amdgpu_buffer_rsrc_t rs, unsigned soff, unsigned voff,
                                       float v)
{
    if constexpr (sizeof(T) == 4)
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), rs, (int)voff,
                                              (int)soff, 0);
    else
        __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, wst<T>(v)), rs,
                                              (int)voff, (int)soff, 0);
}

template <int R, int BY, int RY, int NWM>
constexpr size_t rows_lds_bytes()
{
    return 2u * (size_t)(BY * RY + 2 * R) * (size_t)(64 * NWM + 2 * R) * sizeof(float);
}

template <typename T, int R, int BY, int RY, int NWM, bool IMG, bool TAPER>
__global__ void __launch_bounds__(64 * NWM * BY)
    k_stencil_rows(const T *__restrict__ P, T *__restrict__ PP, const float *__restrict__ V,
                   const Coefs cf, const Grid g, const RowGrid rg, const ImageFuse fz,
                   const TaperFuse tf, const float *__restrict__ gt, int *__restrict__ flag)
{
    constexpr int TY = BY * RY;
    constexpr int LW = 64 * NWM + 2 * R;  // LDS row: z in [-R, 64*NWM + R)
    constexpr int LH = TY + 2 * R;
    constexpr int NH = 2 * R;
    constexpr int NYH = (NH + BY - 1) / BY;
    constexpr int NR = 2 * R + 2;
    constexpr unsigned ES = sizeof(T);          // wavefield element bytes
    constexpr unsigned VS = 4u / sizeof(T);     // float / wavefield offset ratio
    extern __shared__ float smem[];
    const T *__restrict__ PS = static_cast<const T *>(fz.ps);

    int by, bx;
    row_block_coords(rg, by, bx);
    const int z = threadIdx.x;
    const int ty = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int wz0 = __builtin_amdgcn_readfirstlane(threadIdx.x) & ~63;  // wave's first z
    const int y0 = R + by * TY;
    const int xb = R + bx * rg.xchunk;
    const int xe = min(xb + rg.xchunk, g.nx - R);
    const unsigned zc = (unsigned)min(z, g.nz - 1);
    const unsigned zoff = zc * ES;   // wavefield
    const unsigned zoffv = zc * 4u;  // V and the image
    const bool zcomp = (z >= R) && (z < g.nz - R);

    // Taper: only waves holding some z < b, only lanes with z < b.  The
    // table g[b] is copied to LDS behind the two plane buffers, so per-plane
    // factor reads wait on the LDS counter, not behind the prefetches.
    float *tab = smem + 2 * LH * LW;
    if (TAPER) {
        const int tid = threadIdx.y * blockDim.x + threadIdx.x;
        for (int i = tid; i < tf.b; i += blockDim.x * blockDim.y) tab[i] = gt[i];
        __syncthreads();
    }
    // uniform factor lat(i) of a lateral border of n cells
    auto lat = [&](int i, int n) -> float {
        const int j = i < tf.b ? tf.b - 1 - i : (i >= n - tf.b ? i - (n - tf.b) : -1);
        const float v = j >= 0 ? tab[max(j, 0)] : 1.0f;
        return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
    };
    const bool wtap = TAPER && (wz0 < tf.b);
    const bool ztap = wtap && (z < tf.b);
    float fzl = 1.0f;
    if (ztap) fzl = tab[tf.b - 1 - z];

    unsigned rowb[RY];  // row byte offsets of the wavefield in a plane, uniform
    unsigned ymask = 0, rmask = 0;  // bit r: row r is updated / holds receivers
    float fyr[RY];
#pragma unroll
    for (int r = 0; r < RY; ++r) {
        const int y = y0 + ty + BY * r;
        const int yc = min(y, g.ny - 1);
        rowb[r] = (unsigned)yc * (unsigned)g.sy * ES;
        if (y < g.ny - R) ymask |= 1u << r;
        if (IMG) {
            const int dy = y - fz.ry0;
            if (dy >= 0 && (dy % fz.rdy) == 0 && (dy / fz.rdy) < fz.rny) rmask |= 1u << r;
        }
        fyr[r] = 1.0f;
        if (wtap) fyr[r] = lat(yc, g.ny);
    }
    const bool rcvz = IMG && (z == fz.rz);
    const bool imgz = IMG && (z >= fz.z0) && (z < fz.z1);

    unsigned yh_b[NYH];
    int yh_lrow[NYH];
    bool yh_valid[NYH];
    float yh_fy[NYH];
#pragma unroll
    for (int i = 0; i < NYH; ++i) {
        const int h = ty + BY * i;
        yh_valid[i] = h < NH;
        const int gy = min(h < R ? y0 - R + h : y0 + TY + (h - R), g.ny - 1);
        yh_lrow[i] = h < R ? h : R + TY + (h - R);
        yh_b[i] = (unsigned)gy * (unsigned)g.sy * ES;
        yh_fy[i] = 1.0f;
        if (wtap) yh_fy[i] = lat(gy, g.ny);
    }

    float q[NR][RY];
    float n_yh[2][NYH];

    auto plane_fx = [&](int x) -> float { return wtap ? lat(x, g.nx) : 1.0f; };
    // Branch-free taper: the lane factors are exactly 1 where no taper applies
    // (z >= b), so every lane multiplies and the result is unchanged there.
    // One application is one in-place pass of the reference: rounded to T.
    auto tap = [&](float v, float fx, float fy) -> float {
        return wrnd<T>(taper1(v, ztap ? fx : 1.0f, ztap ? fy : 1.0f, fzl));
    };
    // Loads are issued raw; the taper of a prefetched value is applied at the
    // start of the step that first uses it, when the load has long landed
    // (tapering at issue would make each wave wait for its newest load).
    auto load_ring = [&](int x, float(&dst)[RY]) {
        const auto rs = plane_rsrc(P + (size_t)x * (size_t)g.sx);
#pragma unroll
        for (int r = 0; r < RY; ++r) dst[r] = ld_buf<T>(rs, rowb[r], zoff);
    };
    auto taper_ring = [&](int x, float(&dst)[RY]) {
        if (TAPER && wtap) {
            const float fx = plane_fx(x);
#pragma unroll
            for (int r = 0; r < RY; ++r) dst[r] = tap(dst[r], fx, fyr[r]);
        }
    };
    auto fetch_halo = [&](int x, int set) {
        const auto rs = plane_rsrc(P + (size_t)x * (size_t)g.sx);
#pragma unroll
        for (int i = 0; i < NYH; ++i)
            n_yh[set][i] = yh_valid[i] ? ld_buf<T>(rs, yh_b[i], zoff) : 0.f;
    };
    auto taper_halo = [&](int x, int set) {
        if (TAPER && wtap) {
            const float fx = plane_fx(x);
#pragma unroll
            for (int i = 0; i < NYH; ++i) n_yh[set][i] = tap(n_yh[set][i], fx, yh_fy[i]);
        }
    };

#pragma unroll
    for (int s = 0; s <= 2 * R; ++s) load_ring(xb - R + s, q[s]);
    fetch_halo(xb, 0);
#pragma unroll
    for (int s = 0; s < 2 * R; ++s) taper_ring(xb - R + s, q[s]);  // lead: at the first step

    // per-lane LDS element index of this thread's first row, centre column
    const int lbase = (R + ty) * LW + R + z;
    bool bad = false;
    int x = xb;
    auto step = [&]<int K>(std::integral_constant<int, K>) -> bool {
        if (x >= xe) return false;
        constexpr int cur = K & 1;
        constexpr int ctr = (K + R) % NR;
        // prefetched last step (or in the prologue): lead plane x + R, halo of x
        taper_ring(x + R, q[(K + 2 * R) % NR]);
        taper_halo(x, cur);
        float *tl = smem + cur * (LH * LW);
        const size_t pl = (size_t)x * (size_t)g.sx;
        const auto rpp = plane_rsrc(PP + pl);
        const auto rv = plane_rsrc(V + pl);
        const auto rps = plane_rsrc(IMG ? (const void *)(PS + pl) : (const void *)PP);
        const auto rimg = plane_rsrc(IMG ? (const void *)(fz.img + pl) : (const void *)PP);
        // this plane's PP, V (and imaging operands), consumed after the
        // whole plane's Laplacians, which hide their latency
        float c_pp[RY], c_v[RY], c_ps[RY], c_img[RY];
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            c_pp[r] = ld_buf<T>(rpp, rowb[r], zoff);
            c_v[r] = ld_buf<float>(rv, rowb[r] * VS, zoffv);
        }
        if (IMG) {
#pragma unroll
            for (int r = 0; r < RY; ++r) {
                c_ps[r] = ld_buf<T>(rps, rowb[r], zoff);
                c_img[r] = ld_buf<float>(rimg, rowb[r] * VS, zoffv);
            }
        }
#pragma unroll
        for (int r = 0; r < RY; ++r) tl[lbase + BY * r * LW] = q[ctr][r];
#pragma unroll
        for (int i = 0; i < NYH; ++i)
            if (yh_valid[i]) tl[yh_lrow[i] * LW + R + z] = n_yh[cur][i];
        if (x + 1 < xe) {
            load_ring(x + 1 + R, q[(K + 2 * R + 1) % NR]);
            fetch_halo(x + 1, cur ^ 1);
        }
        __syncthreads();

        bool rcvx = false;
        if (IMG) {
            const int dx = x - fz.rx0;
            rcvx = dx >= 0 && (dx % fz.rdx) == 0 && (dx / fz.rdx) < fz.rnx;
        }
        float fxc = 1.0f;
        if (TAPER) fxc = plane_fx(x);
        float lapl[RY];
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            const float *t = tl + lbase + BY * r * LW;
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
            lapl[r] = ax + ay + az;
            // Pin the Laplacian here: left alone, the compiler sinks it into
            // the divergent store branch below and hoists every row's LDS
            // operands above all branches, doubling the register count.
            asm volatile("" : "+v"(lapl[r]));
        }
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            const float p = q[ctr][r];
            float pp = c_pp[r];
            if (TAPER && wtap) {
                pp = tap(pp, fxc, fyr[r]);
                if (tf.pp_taps > 1) pp = tap(pp, fxc, fyr[r]);
            }
            const float npp = 2.f * p - pp + c_v[r] * lapl[r];
            if (((ymask >> r) & 1u) && zcomp) {
                st_buf<T>(rpp, rowb[r], zoff, npp);
                bad |= (npp != npp);
                // the image reads the stored (rounded) receiver value
                if (IMG && imgz && !(rcvx && ((rmask >> r) & 1u) && rcvz))
                    st_buf<float>(rimg, rowb[r] * VS, zoffv, c_img[r] + c_ps[r] * wrnd<T>(npp));
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

// ---------------------------------------------------------------------------
// Barrier-free wave variant.  Every wave is its own tile: 64 z lanes x RY
// rows, marching a chunk of x planes, with no block-wide synchronisation.
//   * x taps: the register ring of 2R+2 planes (as above);
//   * y taps: registers -- the wave's own RY rows of the centre plane plus
//     2R halo rows (R above, R below) loaded a plane ahead;
//   * z taps: the wave's private LDS rows (64 + 2R wide) holding the centre
//     plane and the 2R z-halo cells of each row (one load per lane for
//     RY <= 8).  Only lanes of the same wave touch a region, so the wave's
//     in-order LDS queue plus a wave barrier (a compiler fence, no
//     instruction) orders writes and reads.
// Waves of one block take consecutive z tiles of the same rows, so z-halo
// loads hit lines their neighbours fetch.
// Minimum waves per SIMD asked of the compiler for the barrier-free kernel:
// the register budget (512 / waves) it must fit.
template <typename T, int RY, bool IMG, int TAPER>
constexpr int wave_min_occupancy()
{
    // (every storage type: the tapered 8-row form spills a few bytes at 3
    // waves and is still faster than at 2, FP32 and 16-bit alike)
    return RY == 8 ? (IMG ? 2 : 3) : (IMG || TAPER ? 3 : 4);
}

struct WaveGrid {
    int nzt, nyt, nxc, xchunk;
    long long waves;  // nzt * nyt * nxc
    int zt0;          // first z tile of the launch (its nzt tiles follow)
    int zb;           // two-z waves: z of tile 0 (even); tile t starts at zb + 128 t
};

// One wave's tile of the barrier-free stencil: z tile zt (64 lanes), rows
// y0 = R + yt*RY .. +RY, planes [xb, xe); lt is the wave's private LDS rows,
// tab the taper table in LDS (TAPER).  Shared by the single-grid and the
// batched kernels.
//
// TAPER 1 fuses the upper taper (taperUpper3D), TAPER 2 the taper of every
// border (taper3D): a value is multiplied, in the reference's per-region
// order, by ((v*fy)*fx)*fz on the z border and by (v*fx)*fy off it, factors
// 1 off their border.  taper3D's front-right edge (x < b, y >= ny-b, z
// inside) takes the face-tapered value of (b, y, z): those lanes load plane
// b (a uniform byte offset on their lane offset) and apply
// ((v*fy) rounded *fx)*fy.  With TAPER 2 the wave also stores the tapered
// PP of its z-frame lanes (z < R, z >= nz-R), which the step does not update.
// TAPER 3 applies taper3D to the step's output instead: P is read as stored
// (the previous step stored it tapered), the centre PP is tapered once
// (pp_taps 1; pp_taps 0 when memory already holds it tapered) and
// the new value is tapered before it is stored; the front-right cells and
// the frame are written by separate kernels, so this form needs no plane-b
// loads (the pointwise chain only).  TAPER 4 does the same with the upper
// taper, leaving the planes [b - hold, b) of the output untapered.
template <typename T, int R, int RY, bool IMG, int TAPER>
__device__ __forceinline__ void wave_tile(const T *__restrict__ P, T *__restrict__ PP,
                                          const float *__restrict__ V, const Coefs &cf,
                                          const Grid &g, int zt, int yt, int xb, int xe,
                                          const ImageFuse &fz, const TaperFuse &tf,
                                          const float *tab, float (*lt)[64 + 2 * R],
                                          int *__restrict__ flag)
{
    constexpr int NR = 2 * R + 2;
    constexpr int LW = 64 + 2 * R;
    constexpr int NZH = 2 * R;
    static_assert(NZH * RY <= 64, "one z-halo load per lane");
    constexpr unsigned ES = sizeof(T);
    constexpr unsigned VS = 4u / sizeof(T);
    const T *__restrict__ PS = static_cast<const T *>(fz.ps);
    const int lane = threadIdx.x;
    const int z0 = zt * 64, y0 = R + yt * RY;
    const int z = z0 + lane;
    const unsigned zc = (unsigned)min(z, g.nz - 1);
    const unsigned zoff = zc * ES, zoffv = zc * 4u;
    const bool zcomp = (z >= R) && (z < g.nz - R);

    auto lat = [&](int i, int n) -> float {
        const int j = i < tf.b ? tf.b - 1 - i : (i >= n - tf.b ? i - (n - tf.b) : -1);
        return j >= 0 ? tab[max(j, 0)] : 1.0f;
    };
    constexpr bool ALL = TAPER == 2;
    constexpr bool OUT = TAPER == 3;
    constexpr bool EVERY = ALL || OUT;   // every border (taper3D)
    constexpr bool LOADTAP = TAPER == 1 || TAPER == 2;  // P tapered on load
    constexpr bool OUTU = TAPER == 4;   // the upper taper on the output
    // the wave holds a tapered value: some z < b (halo included, or only
    // its own cells with OUTU), or any (EVERY)
    const bool wtap = TAPER && (EVERY || (OUTU ? z0 < tf.b : z0 - R < tf.b));
    const int zq = (int)zc;
    const bool ztap = wtap && (EVERY ? (zq < tf.b || zq >= g.nz - tf.b) : (z < tf.b));
    float fzl = 1.0f;
    if (ztap) fzl = EVERY ? lat(zq, g.nz) : tab[tf.b - 1 - z];
    const bool zmid = ALL && !ztap;  // front-right lanes (on the rows and planes concerned)
    const bool zframe = EVERY && z < g.nz && !zcomp;

    // own rows (uniform)
    unsigned rowb[RY];
    unsigned ymask = 0, rmask = 0, qmask = 0;
    float fyr[RY];
#pragma unroll
    for (int r = 0; r < RY; ++r) {
        const int y = y0 + r;
        const int yc = min(y, g.ny - 1);
        rowb[r] = (unsigned)yc * (unsigned)g.sy * ES;
        if (y < g.ny - R) ymask |= 1u << r;
        if (ALL && yc >= g.ny - tf.b) qmask |= 1u << r;
        if (IMG) {
            const int dy = y - fz.ry0;
            if (dy >= 0 && (dy % fz.rdy) == 0 && (dy / fz.rdy) < fz.rny) rmask |= 1u << r;
        }
        fyr[r] = 1.0f;
        if (wtap) fyr[r] = __builtin_bit_cast(
            float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, lat(yc, g.ny))));
    }
    // y-halo rows (uniform): h < R above the tile, h >= R below
    unsigned yhb[2 * R];
    unsigned qhmask = 0;
    float yhf[2 * R];
#pragma unroll
    for (int h = 0; h < 2 * R; ++h) {
        const int gy = min(h < R ? y0 - R + h : y0 + RY + (h - R), g.ny - 1);
        yhb[h] = (unsigned)gy * (unsigned)g.sy * ES;
        if (ALL && gy >= g.ny - tf.b) qhmask |= 1u << h;
        yhf[h] = 1.0f;
        if (LOADTAP && wtap) yhf[h] = __builtin_bit_cast(
            float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, lat(gy, g.ny))));
    }
    // this lane's z-halo cell: row zr, column zcol of the LDS row
    const bool zh_on = lane < NZH * RY;
    const int zr = min(lane / NZH, RY - 1), zk = lane % NZH;
    const int zhz = zk < R ? z0 - R + zk : z0 + 64 + (zk - R);
    const int zhzc = min(max(zhz, 0), g.nz - 1);
    const int zcol = zk < R ? zk : 64 + zk;
    const int zhy = min(y0 + zr, g.ny - 1);
    const unsigned zhoff = ((unsigned)zhy * (unsigned)g.sy + (unsigned)zhzc) * ES;
    const bool zhtap =
        LOADTAP && wtap && (ALL ? (zhzc < tf.b || zhzc >= g.nz - tf.b) : zhzc < tf.b);
    float zhfy = 1.0f, zhfz = 1.0f;
    if (zhtap || (ALL && wtap)) zhfy = lat(zhy, g.ny);
    if (zhtap) zhfz = ALL ? lat(zhzc, g.nz) : tab[tf.b - 1 - zhzc];
    const bool zhq = ALL && !zhtap && zhy >= g.ny - tf.b;  // front-right lane (planes x < b)

    float q[NR][RY];
    float yh[2][2 * R], zh[2];

    auto plane_fx = [&](int x) -> float {
        return wtap ? __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(
                                                    __builtin_bit_cast(int, lat(x, g.nx))))
                    : 1.0f;
    };
    // one in-place taper pass of the reference at a value, rounded to the
    // storage type; factors are exactly 1 where no taper applies.  zt: the
    // value is on the tapered z border; q: a front-right value (ALL)
    auto tapv = [&](float v, float fx, float fy, bool zt, float fz, bool q) -> float {
        float a = 1.0f, c = 1.0f, d = 1.0f;
        if (zt) {
            a = fy;
            c = fx;
            d = fz;
        } else if (EVERY) {
            a = q ? fy : fx;
            c = q ? fx : fy;
            d = q ? fy : 1.0f;
        }
        v *= a;
        if (ALL && q) v = wrnd<T>(v);  // the face pass stored it
        v *= c;
        v *= d;
        return wrnd<T>(v);
    };
    auto tap = [&](float v, float fx, float fy) -> float {
        return tapv(v, fx, fy, ztap, fzl, false);
    };
    // byte offset from plane x to plane b (front-right values, x < b)
    auto qdelta = [&](int x) -> unsigned { return (unsigned)(tf.b - x) * (unsigned)g.sx * ES; };
    auto load_ring = [&](int x, float(&dst)[RY]) {
        const auto rs = plane_rsrc(P + (size_t)x * (size_t)g.sx);
        if (ALL && x < tf.b && qmask) {
            const unsigned qoff = zmid ? zoff + qdelta(x) : zoff;
#pragma unroll
            for (int r = 0; r < RY; ++r)
                dst[r] = ld_buf<T>(rs, rowb[r], ((qmask >> r) & 1u) ? qoff : zoff);
        } else {
#pragma unroll
            for (int r = 0; r < RY; ++r) dst[r] = ld_buf<T>(rs, rowb[r], zoff);
        }
    };
    auto taper_ring = [&](int x, float(&dst)[RY]) {
        if (LOADTAP && wtap) {
            const float fx = plane_fx(x);
            const bool qx = ALL && x < tf.b;
#pragma unroll
            for (int r = 0; r < RY; ++r)
                dst[r] = tapv(dst[r], fx, fyr[r], ztap, fzl, qx && ((qmask >> r) & 1u) && zmid);
        }
    };
    auto fetch_halo = [&](int x, int set) {
        const auto rs = plane_rsrc(P + (size_t)x * (size_t)g.sx);
        if (ALL && x < tf.b) {
            const unsigned dq = qdelta(x);
            const unsigned qoff = zmid ? zoff + dq : zoff;
#pragma unroll
            for (int h = 0; h < 2 * R; ++h)
                yh[set][h] = ld_buf<T>(rs, yhb[h], ((qhmask >> h) & 1u) ? qoff : zoff);
            zh[set] = ld_buf<T>(rs, 0u, zhq ? zhoff + dq : zhoff);
        } else {
#pragma unroll
            for (int h = 0; h < 2 * R; ++h) yh[set][h] = ld_buf<T>(rs, yhb[h], zoff);
            zh[set] = ld_buf<T>(rs, 0u, zhoff);
        }
    };
    auto taper_halo = [&](int x, int set) {
        if (LOADTAP && wtap) {
            const float fx = plane_fx(x);
            const bool qx = ALL && x < tf.b;
#pragma unroll
            for (int h = 0; h < 2 * R; ++h)
                yh[set][h] = tapv(yh[set][h], fx, yhf[h], ztap, fzl,
                                  qx && ((qhmask >> h) & 1u) && zmid);
            zh[set] = tapv(zh[set], fx, zhfy, zhtap, zhfz, qx && zhq);
        }
    };

#pragma unroll
    for (int s = 0; s <= 2 * R; ++s) load_ring(xb - R + s, q[s]);
    fetch_halo(xb, 0);
#pragma unroll
    for (int s = 0; s < 2 * R; ++s) taper_ring(xb - R + s, q[s]);

    bool bad = false;
    int x = xb;
    auto step = [&]<int K>(std::integral_constant<int, K>) -> bool {
        if (x >= xe) return false;
        constexpr int cur = K & 1;
        constexpr int ctr = (K + R) % NR;
        taper_ring(x + R, q[(K + 2 * R) % NR]);
        taper_halo(x, cur);
        const size_t pl = (size_t)x * (size_t)g.sx;
        const auto rpp = plane_rsrc(PP + pl);
        const auto rv = plane_rsrc(V + pl);
        const auto rps = plane_rsrc(IMG ? (const void *)(PS + pl) : (const void *)PP);
        const auto rimg = plane_rsrc(IMG ? (const void *)(fz.img + pl) : (const void *)PP);
        float c_pp[RY], c_v[RY], c_ps[RY], c_img[RY];
        // front-right cells read PP at plane b, which this wave updates later
        // in its march (the launcher makes the first x chunk reach b)
        const bool qxc = ALL && x < tf.b;
        const unsigned ppq = qxc && zmid ? zoff + qdelta(x) : zoff;
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            c_pp[r] = ld_buf<T>(rpp, rowb[r], ((qmask >> r) & 1u) ? ppq : zoff);
            c_v[r] = ld_buf<float>(rv, rowb[r] * VS, zoffv);
        }
        if (IMG) {
#pragma unroll
            for (int r = 0; r < RY; ++r) {
                c_ps[r] = ld_buf<T>(rps, rowb[r], zoff);
                c_img[r] = ld_buf<float>(rimg, rowb[r] * VS, zoffv);
            }
        }
        // stage the centre plane's rows and z halo in this wave's LDS rows
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int r = 0; r < RY; ++r) lt[r][R + lane] = q[ctr][r];
        if (zh_on) lt[zr][zcol] = zh[cur];
        __builtin_amdgcn_wave_barrier();
        float yc[2 * R];  // this plane's y-halo values (registers)
#pragma unroll
        for (int h = 0; h < 2 * R; ++h) yc[h] = yh[cur][h];
        if (x + 1 < xe) {
            load_ring(x + 1 + R, q[(K + 2 * R + 1) % NR]);
            fetch_halo(x + 1, cur ^ 1);
        }
        bool rcvx = false;
        if (IMG) {
            const int dx = x - fz.rx0;
            rcvx = dx >= 0 && (dx % fz.rdx) == 0 && (dx / fz.rdx) < fz.rnx;
        }
        float fxc = 1.0f;
        if (TAPER) fxc = plane_fx(x);
        float lapl[RY];
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            const float *t = &lt[r][R + lane];
            const float p = q[ctr][r];
            float ax = 0.f, ay = 0.f, az = 0.f;
#pragma unroll
            for (int io = 0; io <= 2 * R; ++io) {
                const int ry = r + io - R;  // compile-time row of the y tap
                const float vy = ry < 0 ? yc[ry + R] : (ry >= RY ? yc[R + ry - RY] : q[ctr][ry]);
                const float vz = (io == R) ? p : t[io - R];
                ax += q[(K + io) % NR][r] * cf.c[0][io];
                ay += vy * cf.c[1][io];
                az += vz * cf.c[2][io];
            }
            lapl[r] = ax + ay + az;
            asm volatile("" : "+v"(lapl[r]));
        }
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            const float p = q[ctr][r];
            float pp = c_pp[r];
            if (OUT) {
                if (tf.pp_taps > 0) pp = tapv(pp, fxc, fyr[r], ztap, fzl, false);
            } else if (OUTU) {
                if (wtap) pp = tapv(pp, fxc, fyr[r], ztap, fzl, false);
            } else if (TAPER && wtap) {
                const bool qk = qxc && ((qmask >> r) & 1u) && zmid;
                // a front-right PP after two passes: plane b's face value
                // of the first pass, then the second pass's chain
                if (ALL && qk && tf.pp_taps > 1) pp = wrnd<T>(pp * fyr[r]);
                pp = tapv(pp, fxc, fyr[r], ztap, fzl, qk);
                if (tf.pp_taps > 1 && !qk) pp = tap(pp, fxc, fyr[r]);
            }
            const float npp = 2.f * p - pp + c_v[r] * lapl[r];
            // z frame (not updated): PP tapered, and with OUT once more as
            // the next step's P
            if (EVERY && zframe && ((ymask >> r) & 1u))
                st_buf<T>(rpp, rowb[r], zoff,
                          OUT ? tapv(pp, fxc, fyr[r], ztap, fzl, false) : pp);
            if (((ymask >> r) & 1u) && zcomp) {
                // OUT: the value the step stores (rounded), then taper3D
                const bool otap = OUT || (OUTU && ztap && z < tf.b - tf.hold);
                st_buf<T>(rpp, rowb[r], zoff,
                          otap ? tapv(wrnd<T>(npp), fxc, fyr[r], ztap, fzl, false) : npp);
                bad |= (npp != npp);
                if (IMG && (z >= fz.z0) && (z < fz.z1) &&
                    !(rcvx && ((rmask >> r) & 1u) && z == fz.rz))
                    st_buf<float>(rimg, rowb[r] * VS, zoffv, c_img[r] + c_ps[r] * wrnd<T>(npp));
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

// Global access at a wave-uniform base plus a per-lane 32-bit byte offset.
// The offset is made opaque at each use so that loop-invariant code motion
// cannot fold it into per-lane 64-bit pointers (one VGPR pair per row and
// array); the access then selects the scalar-base + 32-bit-offset form.
template <typename U>
__device__ __forceinline__ U ld_at(const char *base, unsigned off)
{
    asm volatile("" : "+v"(off));
    return *reinterpret_cast<const U *>(base + off);
}
template <typename U>
__device__ __forceinline__ void st_at(char *base, unsigned off, U v)
{
    asm volatile("" : "+v"(off));
    *reinterpret_cast<U *>(base + off) = v;
}

template <typename T, int R, int RY, int WPB, bool IMG, int TAPER>
__global__ void __launch_bounds__(64 * WPB, (wave_min_occupancy<T, RY, IMG, TAPER>()))
    k_stencil_wave(const T *__restrict__ P, T *__restrict__ PP, const float *__restrict__ V,
                   const Coefs cf, const Grid g, const WaveGrid wg, const ImageFuse fz,
                   const TaperFuse tf, const float *__restrict__ gt, int *__restrict__ flag)
{
    constexpr int LW = 64 + 2 * R;  // LDS row: z0 - R .. z0 + 63 + R
    __shared__ float tile[WPB][RY][LW];
    extern __shared__ float tab[];  // taper table g[b] (TAPER only)
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.y);
    if (TAPER) {
        const int tid = threadIdx.y * 64 + threadIdx.x;
        for (int i = tid; i < tf.b; i += 64 * WPB) tab[i] = gt[i];
        __syncthreads();  // the only block barrier, before any wave leaves
    }
    // XCD-aware block order, then waves of a block in z-tile order
    long long w;
    {
        const long long nb = (wg.waves + WPB - 1) / WPB;
        const long long h = blockIdx.x;
        const long long q = nb >> 3, rr = nb & 7, k = h & 7, j = h >> 3;
        const long long t = k * q + (k < rr ? k : rr) + j;
        w = t * WPB + wid;
    }
    if (w >= wg.waves) return;
    const int zt = wg.zt0 + (int)(w % wg.nzt);
    const long long w2 = w / wg.nzt;
    const int yt = (int)(w2 % wg.nyt);
    const int xc = (int)(w2 / wg.nyt);
    const int xb = R + xc * wg.xchunk;
    const int xe = min(xb + wg.xchunk, g.nx - R);
    wave_tile<T, R, RY, IMG, TAPER>(P, PP, V, cf, g, zt, yt, xb, xe, fz, tf, tab, tile[wid], flag);
}

// ---------------------------------------------------------------------------
// Barrier-free waves with two z cells per lane.  A wave's tile is 128 z cells
// x RY rows: lane l owns z0 + 2l and z0 + 2l + 1, so every row access moves
// a pair (one 32-bit word of two 16-bit values, or 8 bytes for FP32) and the
// z taps of both cells come from five 8-byte LDS reads.  Tap order, fusions
// and rounding are those of wave_tile.  Needs even R, even nz and even
// strides: a pair is then aligned and either updated whole or not at all.
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

template <typename T>
__device__ __forceinline__ void ld_buf2(__amdgpu_buffer_rsrc_t rs, unsigned soff, unsigned voff,
                                        float (&v)[2])
{
    if constexpr (sizeof(T) == 4) {
        const u32x2 u = __builtin_amdgcn_raw_buffer_load_b64(rs, (int)voff, (int)soff, 0);
        // Copy the elements out first: this compiler's __builtin_bit_cast of
        // a vector-element expression (u.y) reads element 0.
        const unsigned lo = u.x, hi = u.y;
        v[0] = __builtin_bit_cast(float, lo);
        v[1] = __builtin_bit_cast(float, hi);
    } else {
        const unsigned u = __builtin_amdgcn_raw_buffer_load_b32(rs, (int)voff, (int)soff, 0);
        v[0] = wld(__builtin_bit_cast(T, (unsigned short)(u & 0xffffu)));
        v[1] = wld(__builtin_bit_cast(T, (unsigned short)(u >> 16)));
    }
}
template <typename T>
__device__ __forceinline__ void st_buf2(__amdgpu_buffer_rsrc_t rs, unsigned soff, unsigned voff,
                                        float a, float b)
{
    if constexpr (sizeof(T) == 4) {
        u32x2 u;
        u.x = __builtin_bit_cast(unsigned, a);
        u.y = __builtin_bit_cast(unsigned, b);
        __builtin_amdgcn_raw_buffer_store_b64(u, rs, (int)voff, (int)soff, 0);
    } else {
        const unsigned lo = __builtin_bit_cast(unsigned short, wst<T>(a));
        const unsigned hi = __builtin_bit_cast(unsigned short, wst<T>(b));
        __builtin_amdgcn_raw_buffer_store_b32(lo | (hi << 16), rs, (int)voff, (int)soff, 0);
    }
}

template <bool IMG>
constexpr int wave2_min_occupancy()
{
    return IMG ? 2 : 3;
}

// TAPER 1: the upper taper fused on load; TAPER 3: taper3D applied to the
// output, as in wave_tile (pointwise chains on the centre PP, the new value
// and the z frame; the frame's x-y part and the front-right edge are written
// by separate kernels); TAPER 4: the upper taper applied to the output (the
// centre PP and the new value of cells z < b - hold; the frame by
// k_taper_frame).
template <typename T, int R, int RY, bool IMG, int TAPER>
__device__ __forceinline__ void wave_tile2(const T *__restrict__ P, T *__restrict__ PP,
                                           const float *__restrict__ V, const Coefs &cf,
                                           const Grid &g, int z0, int yt, int xb, int xe,
                                           const ImageFuse &fz, const TaperFuse &tf,
                                           const float *tab, float (*lt)[128 + 2 * R],
                                           int *__restrict__ flag)
{
    constexpr int NR = 2 * R + 2;
    constexpr int NZH = 2 * R;
    static_assert(R % 2 == 0, "cell pairs are updated whole");
    static_assert(NZH * RY <= 64, "one z-halo load per lane");
    constexpr unsigned ES = sizeof(T);
    constexpr unsigned VS = 4u / sizeof(T);
    const T *__restrict__ PS = static_cast<const T *>(fz.ps);
    const int lane = threadIdx.x;
    const int y0 = R + yt * RY;
    const int z = z0 + 2 * lane;  // the lane's first cell
    // nz even: nz - 2 starts the last in-grid pair
    const unsigned zc = (unsigned)min(z, g.nz - 2);
    const unsigned zoff = zc * ES, zoffv = zc * 4u;
    // with z, R and nz even this holds for both cells or for neither
    const bool zcomp = (z >= R) && (z < g.nz - R);

    auto lat = [&](int i, int n) -> float {
        const int j = i < tf.b ? tf.b - 1 - i : (i >= n - tf.b ? i - (n - tf.b) : -1);
        return j >= 0 ? tab[max(j, 0)] : 1.0f;
    };
    constexpr bool OUT = TAPER == 3;
    constexpr bool OUTU = TAPER == 4;   // the upper taper on the output
    constexpr bool LOADTAP = TAPER == 1;
    const bool wtap = TAPER && (OUT || (OUTU ? z0 < tf.b : z0 - R < tf.b));
    bool ztap[2];
    float fzl[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const int za = min(z + a, g.nz - 1);
        ztap[a] = wtap && (OUT ? (za < tf.b || za >= g.nz - tf.b) : (z + a < tf.b));
        fzl[a] = 1.0f;
        if (ztap[a]) fzl[a] = OUT ? lat(za, g.nz) : tab[tf.b - 1 - (z + a)];
    }
    const bool zframe = OUT && z < g.nz && !zcomp;

    unsigned rowb[RY];
    unsigned ymask = 0, rmask = 0;
    float fyr[RY];
#pragma unroll
    for (int r = 0; r < RY; ++r) {
        const int y = y0 + r;
        const int yc = min(y, g.ny - 1);
        rowb[r] = (unsigned)yc * (unsigned)g.sy * ES;
        if (y < g.ny - R) ymask |= 1u << r;
        if (IMG) {
            const int dy = y - fz.ry0;
            if (dy >= 0 && (dy % fz.rdy) == 0 && (dy / fz.rdy) < fz.rny) rmask |= 1u << r;
        }
        fyr[r] = 1.0f;
        if (wtap) fyr[r] = __builtin_bit_cast(
            float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, lat(yc, g.ny))));
    }
    unsigned yhb[2 * R];
    float yhf[2 * R];
#pragma unroll
    for (int h = 0; h < 2 * R; ++h) {
        const int gy = min(h < R ? y0 - R + h : y0 + RY + (h - R), g.ny - 1);
        yhb[h] = (unsigned)gy * (unsigned)g.sy * ES;
        yhf[h] = 1.0f;
        if (LOADTAP && wtap) yhf[h] = __builtin_bit_cast(
            float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, lat(gy, g.ny))));
    }
    // this lane's z-halo cell (one 16- or 32-bit load): row zr, LDS column zcol
    const bool zh_on = lane < NZH * RY;
    const int zr = min(lane / NZH, RY - 1), zk = lane % NZH;
    const int zhz = zk < R ? z0 - R + zk : z0 + 128 + (zk - R);
    const int zhzc = min(max(zhz, 0), g.nz - 1);
    const int zcol = zk < R ? zk : 128 + zk;
    const int zhy = min(y0 + zr, g.ny - 1);
    const unsigned zhoff = ((unsigned)zhy * (unsigned)g.sy + (unsigned)zhzc) * ES;
    const bool zhtap = LOADTAP && wtap && zhzc < tf.b;
    float zhfy = 1.0f, zhfz = 1.0f;
    if (zhtap) {
        zhfy = lat(zhy, g.ny);
        zhfz = tab[tf.b - 1 - zhzc];
    }

    float q[NR][RY][2];
    float yh[2][2 * R][2], zh[2];

    auto plane_fx = [&](int x) -> float {
        return wtap ? __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(
                                                    __builtin_bit_cast(int, lat(x, g.nx))))
                    : 1.0f;
    };
    auto tap = [&](float v, int a, float fx, float fy) -> float {
        return wrnd<T>(taper1(v, ztap[a] ? fx : 1.0f, ztap[a] ? fy : 1.0f, fzl[a]));
    };
    // taper3D at a value (OUT): ((v*fy)*fx)*fz on the z border, (v*fx)*fy off
    // it, factors exactly 1 off their border
    auto tapo = [&](float v, int a, float fx, float fy) -> float {
        if (ztap[a]) {
            v *= fy;
            v *= fx;
            v *= fzl[a];
        } else {
            v *= fx;
            v *= fy;
        }
        return wrnd<T>(v);
    };
    auto load_ring = [&](int x, float(&dst)[RY][2]) {
        const auto rs = plane_rsrc(P + (size_t)x * (size_t)g.sx);
#pragma unroll
        for (int r = 0; r < RY; ++r) ld_buf2<T>(rs, rowb[r], zoff, dst[r]);
    };
    auto taper_ring = [&](int x, float(&dst)[RY][2]) {
        if (LOADTAP && wtap) {
            const float fx = plane_fx(x);
#pragma unroll
            for (int r = 0; r < RY; ++r)
#pragma unroll
                for (int a = 0; a < 2; ++a) dst[r][a] = tap(dst[r][a], a, fx, fyr[r]);
        }
    };
    auto fetch_halo = [&](int x, int set) {
        const auto rs = plane_rsrc(P + (size_t)x * (size_t)g.sx);
#pragma unroll
        for (int h = 0; h < 2 * R; ++h) ld_buf2<T>(rs, yhb[h], zoff, yh[set][h]);
        zh[set] = ld_buf<T>(rs, 0u, zhoff);
    };
    auto taper_halo = [&](int x, int set) {
        if (LOADTAP && wtap) {
            const float fx = plane_fx(x);
#pragma unroll
            for (int h = 0; h < 2 * R; ++h)
#pragma unroll
                for (int a = 0; a < 2; ++a) yh[set][h][a] = tap(yh[set][h][a], a, fx, yhf[h]);
            zh[set] = wrnd<T>(taper1(zh[set], zhtap ? fx : 1.0f, zhfy, zhfz));
        }
    };

#pragma unroll
    for (int s = 0; s <= 2 * R; ++s) load_ring(xb - R + s, q[s]);
    fetch_halo(xb, 0);
#pragma unroll
    for (int s = 0; s < 2 * R; ++s) taper_ring(xb - R + s, q[s]);

    bool bad = false;
    int x = xb;
    auto step = [&]<int K>(std::integral_constant<int, K>) -> bool {
        if (x >= xe) return false;
        constexpr int cur = K & 1;
        constexpr int ctr = (K + R) % NR;
        taper_ring(x + R, q[(K + 2 * R) % NR]);
        taper_halo(x, cur);
        const size_t pl = (size_t)x * (size_t)g.sx;
        const auto rpp = plane_rsrc(PP + pl);
        const auto rv = plane_rsrc(V + pl);
        const auto rps = plane_rsrc(IMG ? (const void *)(PS + pl) : (const void *)PP);
        const auto rimg = plane_rsrc(IMG ? (const void *)(fz.img + pl) : (const void *)PP);
        float c_pp[RY][2], c_v[RY][2], c_ps[RY][2], c_img[RY][2];
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            ld_buf2<T>(rpp, rowb[r], zoff, c_pp[r]);
            ld_buf2<float>(rv, rowb[r] * VS, zoffv, c_v[r]);
        }
        if (IMG) {
#pragma unroll
            for (int r = 0; r < RY; ++r) {
                ld_buf2<T>(rps, rowb[r], zoff, c_ps[r]);
                ld_buf2<float>(rimg, rowb[r] * VS, zoffv, c_img[r]);
            }
        }
        // stage the centre plane's rows and z halo in this wave's LDS rows
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int r = 0; r < RY; ++r)
            *reinterpret_cast<float2 *>(&lt[r][R + 2 * lane]) =
                make_float2(q[ctr][r][0], q[ctr][r][1]);
        if (zh_on) lt[zr][zcol] = zh[cur];
        __builtin_amdgcn_wave_barrier();
        float yc[2 * R][2];
#pragma unroll
        for (int h = 0; h < 2 * R; ++h) {
            yc[h][0] = yh[cur][h][0];
            yc[h][1] = yh[cur][h][1];
        }
        if (x + 1 < xe) {
            load_ring(x + 1 + R, q[(K + 2 * R + 1) % NR]);
            fetch_halo(x + 1, cur ^ 1);
        }
        bool rcvx = false;
        if (IMG) {
            const int dx = x - fz.rx0;
            rcvx = dx >= 0 && (dx % fz.rdx) == 0 && (dx / fz.rdx) < fz.rnx;
        }
        float fxc = 1.0f;
        if (TAPER) fxc = plane_fx(x);
        float lapl[RY][2];
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            // LDS row entries 2l .. 2l + 2R + 1: cells z0 - R + 2l .. z0 + R + 2l + 1
            float w[2 * R + 2];
            const float2 *t2 = reinterpret_cast<const float2 *>(&lt[r][2 * lane]);
#pragma unroll
            for (int k = 0; k <= R; ++k) {
                const float2 u = t2[k];
                w[2 * k] = u.x;
                w[2 * k + 1] = u.y;
            }
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                const float p = q[ctr][r][a];
                float ax = 0.f, ay = 0.f, az = 0.f;
#pragma unroll
                for (int io = 0; io <= 2 * R; ++io) {
                    const int ry = r + io - R;  // compile-time row of the y tap
                    const float vy = ry < 0 ? yc[ry + R][a]
                                            : (ry >= RY ? yc[R + ry - RY][a] : q[ctr][ry][a]);
                    const float vz = (io == R) ? p : w[io + a];
                    ax += q[(K + io) % NR][r][a] * cf.c[0][io];
                    ay += vy * cf.c[1][io];
                    az += vz * cf.c[2][io];
                }
                lapl[r][a] = ax + ay + az;
                asm volatile("" : "+v"(lapl[r][a]));
            }
        }
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            float npp[2], ppt[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                float pp = c_pp[r][a];
                if (OUT) {
                    if (tf.pp_taps > 0) pp = tapo(pp, a, fxc, fyr[r]);
                } else if (OUTU) {
                    if (wtap) pp = tap(pp, a, fxc, fyr[r]);
                } else if (TAPER && wtap) {
                    pp = tap(pp, a, fxc, fyr[r]);
                    if (tf.pp_taps > 1) pp = tap(pp, a, fxc, fyr[r]);
                }
                ppt[a] = pp;
                npp[a] = 2.f * q[ctr][r][a] - pp + c_v[r][a] * lapl[r][a];
            }
            // z frame (not updated): PP tapered once more as the next step's P
            if (OUT && zframe && ((ymask >> r) & 1u))
                st_buf2<T>(rpp, rowb[r], zoff, tapo(ppt[0], 0, fxc, fyr[r]),
                           tapo(ppt[1], 1, fxc, fyr[r]));
            if (((ymask >> r) & 1u) && zcomp) {
                if (OUT) {
                    st_buf2<T>(rpp, rowb[r], zoff, tapo(wrnd<T>(npp[0]), 0, fxc, fyr[r]),
                               tapo(wrnd<T>(npp[1]), 1, fxc, fyr[r]));
                } else if (OUTU && wtap) {
                    // cells z < b - hold: the value stored, then the upper taper
                    float o[2];
#pragma unroll
                    for (int a = 0; a < 2; ++a)
                        o[a] = (ztap[a] && z + a < tf.b - tf.hold)
                                   ? tap(wrnd<T>(npp[a]), a, fxc, fyr[r])
                                   : npp[a];
                    st_buf2<T>(rpp, rowb[r], zoff, o[0], o[1]);
                } else {
                    st_buf2<T>(rpp, rowb[r], zoff, npp[0], npp[1]);
                }
                bad |= (npp[0] != npp[0]) || (npp[1] != npp[1]);
                if (IMG) {
                    bool im[2];
                    float iv[2];
#pragma unroll
                    for (int a = 0; a < 2; ++a) {
                        const int za = z + a;
                        im[a] = za >= fz.z0 && za < fz.z1 &&
                                !(rcvx && ((rmask >> r) & 1u) && za == fz.rz);
                        iv[a] = c_img[r][a] + c_ps[r][a] * wrnd<T>(npp[a]);
                    }
                    if (im[0] && im[1]) st_buf2<float>(rimg, rowb[r] * VS, zoffv, iv[0], iv[1]);
                    else if (im[0]) st_buf<float>(rimg, rowb[r] * VS, zoffv, iv[0]);
                    else if (im[1]) st_buf<float>(rimg, rowb[r] * VS, zoffv + 4u, iv[1]);
                }
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

template <typename T, int R, int RY, int WPB, bool IMG, int TAPER>
__global__ void __launch_bounds__(64 * WPB, (wave2_min_occupancy<IMG>()))
    k_stencil_wave2(const T *__restrict__ P, T *__restrict__ PP, const float *__restrict__ V,
                    const Coefs cf, const Grid g, const WaveGrid wg, const ImageFuse fz,
                    const TaperFuse tf, const float *__restrict__ gt, int *__restrict__ flag)
{
    constexpr int LW = 128 + 2 * R;  // LDS row: z0 - R .. z0 + 127 + R
    __shared__ __align__(8) float tile[WPB][RY][LW];
    extern __shared__ float tab[];
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.y);
    if (TAPER) {
        const int tid = threadIdx.y * 64 + threadIdx.x;
        for (int i = tid; i < tf.b; i += 64 * WPB) tab[i] = gt[i];
        __syncthreads();
    }
    long long w;
    {
        const long long nb = (wg.waves + WPB - 1) / WPB;
        const long long h = blockIdx.x;
        const long long q = nb >> 3, rr = nb & 7, k = h & 7, j = h >> 3;
        const long long t = k * q + (k < rr ? k : rr) + j;
        w = t * WPB + wid;
    }
    if (w >= wg.waves) return;
    const int z0 = wg.zb + (wg.zt0 + (int)(w % wg.nzt)) * 128;
    const long long w2 = w / wg.nzt;
    const int yt = (int)(w2 % wg.nyt);
    const int xc = (int)(w2 / wg.nyt);
    const int xb = R + xc * wg.xchunk;
    const int xe = min(xb + wg.xchunk, g.nx - R);
    wave_tile2<T, R, RY, IMG, TAPER>(P, PP, V, cf, g, z0, yt, xb, xe, fz, tf, tab, tile[wid],
                                     flag);
}

// Batched form: the waves of every shot of a batch in one launch.  Each
// shot's waves fill whole blocks (blocks never straddle shots), so a block
// finds its shot by a binary search over the shots' first blocks and keeps
// one taper table.  The block order is XCD-aware as above, so each XCD
// marches a contiguous run of tiles.
template <typename T, int R, int RY, int WPB, bool IMG, bool TAPER>
__global__ void __launch_bounds__(64 * WPB, (wave_min_occupancy<T, RY, IMG, TAPER>()))
    k_stencil_batch(const BatchShot *__restrict__ shots, int nshots, long long nblocks, int par,
                    int pspar, const Coefs cf, const TaperFuse tf, int *__restrict__ flag)
{
    constexpr int LW = 64 + 2 * R;
    __shared__ float tile[WPB][RY][LW];
    extern __shared__ float tab[];
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.y);
    if (TAPER) {
        const int tid = threadIdx.y * 64 + threadIdx.x;
        for (int i = tid; i < tf.b; i += 64 * WPB) tab[i] = tf.t[i];
        __syncthreads();
    }
    long long t;
    {
        const long long h = blockIdx.x;
        const long long q = nblocks >> 3, rr = nblocks & 7, k = h & 7, j = h >> 3;
        t = k * q + (k < rr ? k : rr) + j;
    }
    int lo = 0, hi = nshots - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (shots[mid].blk0 <= t) lo = mid;
        else hi = mid - 1;
    }
    const BatchShot &sh = shots[lo];
    const long long w = (t - sh.blk0) * WPB + wid;
    if (w >= sh.nwaves) return;
    const int zt = (int)(w % sh.nzt);
    const long long w2 = w / sh.nzt;
    const int yt = (int)(w2 % sh.nyt);
    const int xc = (int)(w2 / sh.nyt);
    const int xb = R + xc * sh.xchunk;
    const int xe = min(xb + sh.xchunk, sh.g.nx - R);
    void *const *f = IMG ? sh.fr : sh.fs;
    const T *P = static_cast<const T *>(f[par]);
    T *PP = static_cast<T *>(f[par ^ 1]);
    ImageFuse fz{};
    if (IMG) {
        fz.img = sh.img;
        fz.ps = sh.fs[pspar];
        fz.z0 = tf.b;
        fz.z1 = sh.g.nz - tf.b;
        fz.rz = sh.rz;
        fz.rx0 = sh.rx0;
        fz.rdx = sh.rdx;
        fz.rnx = sh.rnx;
        fz.ry0 = sh.ry0;
        fz.rdy = sh.rdy;
        fz.rny = sh.rny;
    }
    wave_tile<T, R, RY, IMG, TAPER>(P, PP, sh.V, cf, sh.g, zt, yt, xb, xe, fz, tf, tab,
                                    tile[wid], flag);
}

// ---------------------------------------------------------------------------
// Full-row variant with two z per lane.  A lane owns z = 2*tx and 2*tx + 1,
// so every access moves a pair: 8 bytes for FP32 wavefields, one 32-bit word
// of two 16-bit values otherwise (one 2-byte value per lane streams at well
// under half the 4-byte rate; 8-byte lanes stream about 10 % faster than
// 4-byte ones).  The block layout is the full-row one (BY thread rows x RY
// rows, 64 * NW lanes per row, NW = ceil(nz / 128)); needs nz even (pairs
// never straddle a row end).  Accesses are global loads at a wave-uniform
// row base plus the lane's 32-bit byte offset (see ld_at).
template <typename T, int R, int BY, int RY, int NWM, bool IMG, bool TAPER>
__global__ void __launch_bounds__(64 * NWM * BY)
    k_stencil_rows2(const T *__restrict__ P, T *__restrict__ PP, const float *__restrict__ V,
                    const Coefs cf, const Grid g, const RowGrid rg, const ImageFuse fz,
                    const TaperFuse tf, const float *__restrict__ gt, int *__restrict__ flag)
{
    static_assert(R % 2 == 0, "8-byte aligned LDS pairs");
    constexpr int TY = BY * RY;
    constexpr int LW = 128 * NWM + 2 * R;  // LDS row: z in [-R, 128*NWM + R)
    constexpr int LH = TY + 2 * R;
    constexpr int NH = 2 * R;
    constexpr int NYH = (NH + BY - 1) / BY;
    constexpr int NR = 2 * R + 2;
    extern __shared__ float smem[];
    const T *__restrict__ PS = static_cast<const T *>(fz.ps);

    int by, bx;
    row_block_coords(rg, by, bx);
    const int tx = threadIdx.x;
    const int ty = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int wz0 = 2 * (__builtin_amdgcn_readfirstlane(threadIdx.x) & ~63);
    const int y0 = R + by * TY;
    const int xb = R + bx * rg.xchunk;
    const int xe = min(xb + rg.xchunk, g.nx - R);
    const int zl = 2 * tx;                          // first z of the pair
    const unsigned zp = (unsigned)min(zl, g.nz - 2);  // clamped pair start
    constexpr unsigned ES = sizeof(T);
    constexpr unsigned VS = 4u / sizeof(T);         // float / wavefield offset ratio
    const unsigned zoff = zp * ES;                  // wavefield byte offset
    const unsigned zoffv = zp * 4u;                 // float byte offset
    bool zcomp[2];
#pragma unroll
    for (int e = 0; e < 2; ++e) zcomp[e] = (zl + e >= R) && (zl + e < g.nz - R);

    float *tab = smem + 2 * LH * LW;
    if (TAPER) {
        const int tid = threadIdx.y * blockDim.x + threadIdx.x;
        for (int i = tid; i < tf.b; i += blockDim.x * blockDim.y) tab[i] = gt[i];
        __syncthreads();
    }
    auto lat = [&](int i, int n) -> float {
        const int j = i < tf.b ? tf.b - 1 - i : (i >= n - tf.b ? i - (n - tf.b) : -1);
        const float v = j >= 0 ? tab[max(j, 0)] : 1.0f;
        return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
    };
    const bool wtap = TAPER && (wz0 < tf.b);
    bool ztap[2];
    float fzl[2];
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        ztap[e] = wtap && (zl + e < tf.b);
        fzl[e] = ztap[e] ? tab[tf.b - 1 - (zl + e)] : 1.0f;
    }

    unsigned rowb[RY];  // wavefield row byte offsets (uniform)
    unsigned ymask = 0, rmask = 0;
    float fyr[RY];
#pragma unroll
    for (int r = 0; r < RY; ++r) {
        const int y = y0 + ty + BY * r;
        const int yc = min(y, g.ny - 1);
        rowb[r] = (unsigned)yc * (unsigned)g.sy * ES;
        if (y < g.ny - R) ymask |= 1u << r;
        if (IMG) {
            const int dy = y - fz.ry0;
            if (dy >= 0 && (dy % fz.rdy) == 0 && (dy / fz.rdy) < fz.rny) rmask |= 1u << r;
        }
        fyr[r] = 1.0f;
        if (wtap) fyr[r] = lat(yc, g.ny);
    }
    unsigned yh_b[NYH];
    int yh_lrow[NYH];
    bool yh_valid[NYH];
    float yh_fy[NYH];
#pragma unroll
    for (int i = 0; i < NYH; ++i) {
        const int h = ty + BY * i;
        yh_valid[i] = h < NH;
        const int gy = min(h < R ? y0 - R + h : y0 + TY + (h - R), g.ny - 1);
        yh_lrow[i] = h < R ? h : R + TY + (h - R);
        yh_b[i] = (unsigned)gy * (unsigned)g.sy * ES;
        yh_fy[i] = 1.0f;
        if (wtap) yh_fy[i] = lat(gy, g.ny);
    }

    // pair load / store of the wavefield (widened to / rounded from float)
    // and of float arrays, at plane + row byte offset + the lane's offset
    auto ld2 = [&](const T *plane, unsigned soff, float(&d)[2]) {
        const char *base = reinterpret_cast<const char *>(plane) + soff;
        if constexpr (sizeof(T) == 4) {
            const float2 v = ld_at<float2>(base, zoff);
            d[0] = v.x;
            d[1] = v.y;
        } else {
            const unsigned w = ld_at<unsigned>(base, zoff);
            d[0] = wld(__builtin_bit_cast(T, (unsigned short)(w & 0xffffu)));
            d[1] = wld(__builtin_bit_cast(T, (unsigned short)(w >> 16)));
        }
    };
    auto ldf2 = [&](const float *plane, unsigned soff, float(&d)[2]) {
        const float2 v = ld_at<float2>(reinterpret_cast<const char *>(plane) + soff, zoffv);
        d[0] = v.x;
        d[1] = v.y;
    };

    float q[NR][RY][2];
    float n_yh[2][NYH][2];

    auto plane_fx = [&](int x) -> float { return wtap ? lat(x, g.nx) : 1.0f; };
    auto tap = [&](float v, int e, float fx, float fy) -> float {
        return wrnd<T>(taper1(v, ztap[e] ? fx : 1.0f, ztap[e] ? fy : 1.0f, fzl[e]));
    };
    auto load_ring = [&](int x, float(&dst)[RY][2]) {
        const T *pl = P + (size_t)x * (size_t)g.sx;
#pragma unroll
        for (int r = 0; r < RY; ++r) ld2(pl, rowb[r], dst[r]);
    };
    auto taper_ring = [&](int x, float(&dst)[RY][2]) {
        if (TAPER && wtap) {
            const float fx = plane_fx(x);
#pragma unroll
            for (int r = 0; r < RY; ++r)
#pragma unroll
                for (int e = 0; e < 2; ++e) dst[r][e] = tap(dst[r][e], e, fx, fyr[r]);
        }
    };
    auto fetch_halo = [&](int x, int set) {
        const T *pl = P + (size_t)x * (size_t)g.sx;
#pragma unroll
        for (int i = 0; i < NYH; ++i) {
            if (yh_valid[i]) {
                ld2(pl, yh_b[i], n_yh[set][i]);
            } else {
                n_yh[set][i][0] = 0.f;
                n_yh[set][i][1] = 0.f;
            }
        }
    };
    auto taper_halo = [&](int x, int set) {
        if (TAPER && wtap) {
            const float fx = plane_fx(x);
#pragma unroll
            for (int i = 0; i < NYH; ++i)
#pragma unroll
                for (int e = 0; e < 2; ++e)
                    n_yh[set][i][e] = tap(n_yh[set][i][e], e, fx, yh_fy[i]);
        }
    };

#pragma unroll
    for (int s = 0; s <= 2 * R; ++s) load_ring(xb - R + s, q[s]);
    fetch_halo(xb, 0);
#pragma unroll
    for (int s = 0; s < 2 * R; ++s) taper_ring(xb - R + s, q[s]);

    const int lbase = (R + ty) * LW + R + zl;  // first element of the pair, first row
    bool bad = false;
    int x = xb;
    auto step = [&]<int K>(std::integral_constant<int, K>) -> bool {
        if (x >= xe) return false;
        constexpr int cur = K & 1;
        constexpr int ctr = (K + R) % NR;
        taper_ring(x + R, q[(K + 2 * R) % NR]);
        taper_halo(x, cur);
        float *tl = smem + cur * (LH * LW);
        const size_t pl = (size_t)x * (size_t)g.sx;
        T *const ppl = PP + pl;
        const float *const vpl = V + pl;
        float c_pp[RY][2], c_v[RY][2], c_ps[RY][2], c_img[RY][2];
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            ld2(ppl, rowb[r], c_pp[r]);
            ldf2(vpl, rowb[r] * VS, c_v[r]);
        }
        if (IMG) {
#pragma unroll
            for (int r = 0; r < RY; ++r) {
                ld2(PS + pl, rowb[r], c_ps[r]);
                ldf2(fz.img + pl, rowb[r] * VS, c_img[r]);
            }
        }
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            float2 v2v;
            v2v.x = q[ctr][r][0];
            v2v.y = q[ctr][r][1];
            *reinterpret_cast<float2 *>(&tl[lbase + BY * r * LW]) = v2v;
        }
#pragma unroll
        for (int i = 0; i < NYH; ++i)
            if (yh_valid[i]) {
                float2 v2v;
                v2v.x = n_yh[cur][i][0];
                v2v.y = n_yh[cur][i][1];
                *reinterpret_cast<float2 *>(&tl[yh_lrow[i] * LW + R + zl]) = v2v;
            }
        if (x + 1 < xe) {
            load_ring(x + 1 + R, q[(K + 2 * R + 1) % NR]);
            fetch_halo(x + 1, cur ^ 1);
        }
        __syncthreads();

        bool rcvx = false;
        if (IMG) {
            const int dx = x - fz.rx0;
            rcvx = dx >= 0 && (dx % fz.rdx) == 0 && (dx / fz.rdx) < fz.rnx;
        }
        float fxc = 1.0f;
        if (TAPER) fxc = plane_fx(x);
        float lapl[RY][2];
#pragma unroll
        for (int r = 0; r < RY; ++r)
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const float *t = tl + lbase + BY * r * LW + e;
                const float p = q[ctr][r][e];
                float ax = 0.f, ay = 0.f, az = 0.f;
#pragma unroll
                for (int io = 0; io <= 2 * R; ++io) {
                    const float vy = (io == R) ? p : t[(io - R) * LW];
                    const float vz = (io == R) ? p : t[io - R];
                    ax += q[(K + io) % NR][r][e] * cf.c[0][io];
                    ay += vy * cf.c[1][io];
                    az += vz * cf.c[2][io];
                }
                lapl[r][e] = ax + ay + az;
                asm volatile("" : "+v"(lapl[r][e]));
            }
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            float npp[2];
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                float pp = c_pp[r][e];
                if (TAPER && wtap) {
                    pp = tap(pp, e, fxc, fyr[r]);
                    if (tf.pp_taps > 1) pp = tap(pp, e, fxc, fyr[r]);
                }
                npp[e] = 2.f * q[ctr][r][e] - pp + c_v[r][e] * lapl[r][e];
            }
            if ((ymask >> r) & 1u) {
                char *base = reinterpret_cast<char *>(ppl) + rowb[r];
                if constexpr (sizeof(T) == 4) {
                    if (zcomp[0] && zcomp[1]) st_at<float2>(base, zoff, make_float2(npp[0], npp[1]));
                    else if (zcomp[0]) st_at<float>(base, zoff, npp[0]);
                    else if (zcomp[1]) st_at<float>(base, zoff + 4u, npp[1]);
                } else {
                    const unsigned short b0 = __builtin_bit_cast(unsigned short, wst<T>(npp[0]));
                    const unsigned short b1 = __builtin_bit_cast(unsigned short, wst<T>(npp[1]));
                    if (zcomp[0] && zcomp[1])
                        st_at<unsigned>(base, zoff, (unsigned)b0 | ((unsigned)b1 << 16));
                    else if (zcomp[0]) st_at<unsigned short>(base, zoff, b0);
                    else if (zcomp[1]) st_at<unsigned short>(base, zoff + 2u, b1);
                }
#pragma unroll
                for (int e = 0; e < 2; ++e) {
                    if (!zcomp[e]) continue;
                    bad |= (npp[e] != npp[e]);
                    const int ze = zl + e;
                    if (IMG && ze >= fz.z0 && ze < fz.z1 &&
                        !(rcvx && ((rmask >> r) & 1u) && ze == fz.rz))
                        st_at<float>(reinterpret_cast<char *>(fz.img + pl) + rowb[r] * VS,
                                     zoffv + 4u * e, c_img[r][e] + c_ps[r][e] * wrnd<T>(npp[e]));
                }
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

// ---------------------------------------------------------------------------
// Backward step of both wavefields in one pass (barrier-free waves).  At
// time it the reference steps the source field back (RTMHBCKernel backward:
// step, UPB restore, swap), then tapers and steps the receiver field,
// injects the receivers and images img += Ps * PPr on z in [blen,
// nze-blen).  Both updates march the same x planes here: V is read once and
// the image takes the new source value from registers instead of reading it
// back (36 instead of 44 bytes per cell).  The UPB restore rewrites only
// z < blen, which the image never reads, so it runs after this kernel.
// Source: plain step.  Receiver: TAPER (HBC) as in the tapered variants.
// Receiver cells are imaged by the receiver kernel after injection.
template <typename T, int R, int RY, int WPB, bool TAPER>
__global__ void __launch_bounds__(64 * WPB)
    k_stencil_dual(const T *__restrict__ Ps, T *__restrict__ PPs, const T *__restrict__ Pr,
                   T *__restrict__ PPr, const float *__restrict__ V, const Coefs cf,
                   const Grid g, const WaveGrid wg, const ImageFuse fz, const TaperFuse tf,
                   const float *__restrict__ gt, int *__restrict__ flag)
{
    constexpr int NR = 2 * R + 2;
    constexpr int LW = 64 + 2 * R;
    constexpr int NZH = 2 * R;
    static_assert(NZH * RY <= 64, "one z-halo load per lane and field");
    constexpr unsigned ES = sizeof(T);
    constexpr unsigned VS = 4u / sizeof(T);
    __shared__ float tile[WPB][2][RY][LW];  // [wave][field][row][z]
    extern __shared__ float tab[];

    const int lane = threadIdx.x;
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.y);
    if (TAPER) {
        const int tid = threadIdx.y * 64 + lane;
        for (int i = tid; i < tf.b; i += 64 * WPB) tab[i] = gt[i];
        __syncthreads();
    }
    long long w;
    {
        const long long nb = (wg.waves + WPB - 1) / WPB;
        const long long h = blockIdx.x;
        const long long q = nb >> 3, rr = nb & 7, k = h & 7, j = h >> 3;
        const long long t = k * q + (k < rr ? k : rr) + j;
        w = t * WPB + wid;
    }
    if (w >= wg.waves) return;
    const int zt = wg.zt0 + (int)(w % wg.nzt);
    const long long w2 = w / wg.nzt;
    const int yt = (int)(w2 % wg.nyt);
    const int xc = (int)(w2 / wg.nyt);
    const int z0 = zt * 64, y0 = R + yt * RY;
    const int xb = R + xc * wg.xchunk;
    const int xe = min(xb + wg.xchunk, g.nx - R);
    const int z = z0 + lane;
    const unsigned zc = (unsigned)min(z, g.nz - 1);
    const unsigned zoff = zc * ES, zoffv = zc * 4u;
    const bool zcomp = (z >= R) && (z < g.nz - R);
    const bool imgz = (z >= fz.z0) && (z < fz.z1);

    auto lat = [&](int i, int n) -> float {
        const int j = i < tf.b ? tf.b - 1 - i : (i >= n - tf.b ? i - (n - tf.b) : -1);
        return j >= 0 ? tab[max(j, 0)] : 1.0f;
    };
    const bool wtap = TAPER && (z0 - R < tf.b);
    const bool ztap = wtap && (z < tf.b);
    float fzl = 1.0f;
    if (ztap) fzl = tab[tf.b - 1 - z];

    unsigned rowb[RY];
    unsigned ymask = 0, rmask = 0;
    float fyr[RY];
#pragma unroll
    for (int r = 0; r < RY; ++r) {
        const int y = y0 + r;
        const int yc = min(y, g.ny - 1);
        rowb[r] = (unsigned)yc * (unsigned)g.sy * ES;
        if (y < g.ny - R) ymask |= 1u << r;
        const int dy = y - fz.ry0;
        if (dy >= 0 && (dy % fz.rdy) == 0 && (dy / fz.rdy) < fz.rny) rmask |= 1u << r;
        fyr[r] = 1.0f;
        if (wtap) fyr[r] = __builtin_bit_cast(
            float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, lat(yc, g.ny))));
    }
    unsigned yhb[2 * R];
    float yhf[2 * R];
#pragma unroll
    for (int h = 0; h < 2 * R; ++h) {
        const int gy = min(h < R ? y0 - R + h : y0 + RY + (h - R), g.ny - 1);
        yhb[h] = (unsigned)gy * (unsigned)g.sy * ES;
        yhf[h] = 1.0f;
        if (wtap) yhf[h] = __builtin_bit_cast(
            float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, lat(gy, g.ny))));
    }
    const bool zh_on = lane < NZH * RY;
    const int zr = min(lane / NZH, RY - 1), zk = lane % NZH;
    const int zhz = zk < R ? z0 - R + zk : z0 + 64 + (zk - R);
    const int zhzc = min(max(zhz, 0), g.nz - 1);
    const int zcol = zk < R ? zk : 64 + zk;
    const int zhy = min(y0 + zr, g.ny - 1);
    const unsigned zhoff = ((unsigned)zhy * (unsigned)g.sy + (unsigned)zhzc) * ES;
    const bool zhtap = wtap && zhzc < tf.b;
    float zhfy = 1.0f, zhfz = 1.0f;
    if (zhtap) {
        zhfy = lat(zhy, g.ny);
        zhfz = tab[tf.b - 1 - zhzc];
    }

    // field 0: source (plain), field 1: receiver (tapered when TAPER)
    float qs[NR][RY], qr[NR][RY];
    float ys[2][2 * R], yr[2][2 * R], zs[2], zr2[2];

    auto plane_fx = [&](int x) -> float {
        return wtap ? __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(
                                                    __builtin_bit_cast(int, lat(x, g.nx))))
                    : 1.0f;
    };
    auto tap = [&](float v, float fx, float fy) -> float {
        return wrnd<T>(taper1(v, ztap ? fx : 1.0f, ztap ? fy : 1.0f, fzl));
    };
    auto load_ring = [&](const T *F, int x, float(&dst)[RY]) {
        const auto rs = plane_rsrc(F + (size_t)x * (size_t)g.sx);
#pragma unroll
        for (int r = 0; r < RY; ++r) dst[r] = ld_buf<T>(rs, rowb[r], zoff);
    };
    auto taper_ring = [&](int x, float(&dst)[RY]) {
        if (TAPER && wtap) {
            const float fx = plane_fx(x);
#pragma unroll
            for (int r = 0; r < RY; ++r) dst[r] = tap(dst[r], fx, fyr[r]);
        }
    };
    auto fetch_halo = [&](const T *F, int x, float(&yh)[2 * R], float &zh) {
        const auto rs = plane_rsrc(F + (size_t)x * (size_t)g.sx);
#pragma unroll
        for (int h = 0; h < 2 * R; ++h) yh[h] = ld_buf<T>(rs, yhb[h], zoff);
        zh = ld_buf<T>(rs, 0u, zhoff);
    };
    auto taper_halo = [&](int x, float(&yh)[2 * R], float &zh) {
        if (TAPER && wtap) {
            const float fx = plane_fx(x);
#pragma unroll
            for (int h = 0; h < 2 * R; ++h) yh[h] = tap(yh[h], fx, yhf[h]);
            zh = wrnd<T>(taper1(zh, zhtap ? fx : 1.0f, zhfy, zhfz));
        }
    };

#pragma unroll
    for (int s = 0; s <= 2 * R; ++s) {
        load_ring(Ps, xb - R + s, qs[s]);
        load_ring(Pr, xb - R + s, qr[s]);
    }
    fetch_halo(Ps, xb, ys[0], zs[0]);
    fetch_halo(Pr, xb, yr[0], zr2[0]);
#pragma unroll
    for (int s = 0; s < 2 * R; ++s) taper_ring(xb - R + s, qr[s]);

    float(*const lts)[LW] = tile[wid][0];
    float(*const ltr)[LW] = tile[wid][1];
    bool bad = false;
    int x = xb;
    auto step = [&]<int K>(std::integral_constant<int, K>) -> bool {
        if (x >= xe) return false;
        constexpr int cur = K & 1;
        constexpr int ctr = (K + R) % NR;
        taper_ring(x + R, qr[(K + 2 * R) % NR]);
        taper_halo(x, yr[cur], zr2[cur]);
        const size_t pl = (size_t)x * (size_t)g.sx;
        const auto rpps = plane_rsrc(PPs + pl);
        const auto rppr = plane_rsrc(PPr + pl);
        const auto rv = plane_rsrc(V + pl);
        const auto rimg = plane_rsrc(fz.img + pl);
        float c_pps[RY], c_ppr[RY], c_v[RY], c_img[RY];
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            c_pps[r] = ld_buf<T>(rpps, rowb[r], zoff);
            c_ppr[r] = ld_buf<T>(rppr, rowb[r], zoff);
            c_v[r] = ld_buf<float>(rv, rowb[r] * VS, zoffv);
            c_img[r] = ld_buf<float>(rimg, rowb[r] * VS, zoffv);
        }
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            lts[r][R + lane] = qs[ctr][r];
            ltr[r][R + lane] = qr[ctr][r];
        }
        if (zh_on) {
            lts[zr][zcol] = zs[cur];
            ltr[zr][zcol] = zr2[cur];
        }
        __builtin_amdgcn_wave_barrier();
        float ycs[2 * R], ycr[2 * R];
#pragma unroll
        for (int h = 0; h < 2 * R; ++h) {
            ycs[h] = ys[cur][h];
            ycr[h] = yr[cur][h];
        }
        if (x + 1 < xe) {
            load_ring(Ps, x + 1 + R, qs[(K + 2 * R + 1) % NR]);
            load_ring(Pr, x + 1 + R, qr[(K + 2 * R + 1) % NR]);
            fetch_halo(Ps, x + 1, ys[cur ^ 1], zs[cur ^ 1]);
            fetch_halo(Pr, x + 1, yr[cur ^ 1], zr2[cur ^ 1]);
        }
        bool rcvx = false;
        {
            const int dx = x - fz.rx0;
            rcvx = dx >= 0 && (dx % fz.rdx) == 0 && (dx / fz.rdx) < fz.rnx;
        }
        float fxc = 1.0f;
        if (TAPER) fxc = plane_fx(x);
        float lapls[RY], laplr[RY];
#pragma unroll
        for (int r = 0; r < RY; ++r) {
#pragma unroll
            for (int f = 0; f < 2; ++f) {
                const float(&q)[NR][RY] = f ? qr : qs;
                const float(&yc)[2 * R] = f ? ycr : ycs;
                const float *t = f ? &ltr[r][R + lane] : &lts[r][R + lane];
                const float p = q[ctr][r];
                float ax = 0.f, ay = 0.f, az = 0.f;
#pragma unroll
                for (int io = 0; io <= 2 * R; ++io) {
                    const int ry = r + io - R;
                    const float vy =
                        ry < 0 ? yc[ry + R] : (ry >= RY ? yc[R + ry - RY] : q[ctr][ry]);
                    const float vz = (io == R) ? p : t[io - R];
                    ax += q[(K + io) % NR][r] * cf.c[0][io];
                    ay += vy * cf.c[1][io];
                    az += vz * cf.c[2][io];
                }
                float l = ax + ay + az;
                asm volatile("" : "+v"(l));
                if (f) laplr[r] = l;
                else lapls[r] = l;
            }
        }
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            const float nps = 2.f * qs[ctr][r] - c_pps[r] + c_v[r] * lapls[r];
            float pp = c_ppr[r];
            if (TAPER && wtap) {
                pp = tap(pp, fxc, fyr[r]);
                if (tf.pp_taps > 1) pp = tap(pp, fxc, fyr[r]);
            }
            const float npr = 2.f * qr[ctr][r] - pp + c_v[r] * laplr[r];
            if (((ymask >> r) & 1u) && zcomp) {
                st_buf<T>(rpps, rowb[r], zoff, nps);
                st_buf<T>(rppr, rowb[r], zoff, npr);
                bad |= (nps != nps) | (npr != npr);
                if (imgz && !(rcvx && ((rmask >> r) & 1u) && z == fz.rz))
                    st_buf<float>(rimg, rowb[r] * VS, zoffv,
                                  c_img[r] + wrnd<T>(nps) * wrnd<T>(npr));
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
    size_t lds;
    bool operator<(const LaunchKey &o) const
    {
        if (fn != o.fn) return fn < o.fn;
        if (threads != o.threads) return threads < o.threads;
        return lds < o.lds;
    }
};

// Depth of active stream captures (capture_guard).  While it is non-zero,
// device_capacity answers from its cache only: hipFuncSetAttribute and the
// occupancy query are never issued inside a capture.
std::atomic<int> g_capture_depth{0};

// Resident blocks on the whole device for one kernel (occupancy x CUs).
// Kernels using more than 64 KiB of dynamic LDS get the attribute raised
// first (gfx950 has 160 KiB per CU).  Returns 0 for a launch shape first seen
// during a capture; the launchers then fail with
// hipErrorStreamCaptureUnsupported rather than query the runtime there.
int device_capacity(const void *fn, int threads, size_t lds = 0)
{
    static std::mutex mu;
    static std::map<LaunchKey, int> cache;
    std::lock_guard<std::mutex> lock(mu);
    const LaunchKey key{fn, threads, lds};
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    if (g_capture_depth.load() > 0) return 0;
    if (lds > 65536)
        (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    int dev = 0, ncu = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        ncu = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, lds) != hipSuccess ||
        per_cu <= 0)
        per_cu = 1;
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

template <typename T, int R, int TZ, int BY, int RY, bool IMG, bool TAPER>
hipError_t launch_cfg(const T *P, T *PP, const float *V, const Coefs &cf, const Grid &g,
                      const ImageFuse &fz, const TaperFuse &tf, int *flag, hipStream_t s,
                      int xchunk)
{
    constexpr int TY = BY * RY;
    auto kern = k_stencil<T, R, TZ, BY, RY, IMG, TAPER>;
    TileGrid tg;
    tg.ntz = (g.nz + TZ - 1) / TZ;
    tg.nty = (g.ny - 2 * R + TY - 1) / TY;
    const int nxi = g.nx - 2 * R;
    if (tg.nty <= 0 || nxi <= 0 || g.nz - 2 * R <= 0) return hipSuccess;  // nothing to update
    if (xchunk <= 0) {
        const int cap = device_capacity((const void *)kern, TZ * BY);
        if (cap <= 0) return hipErrorStreamCaptureUnsupported;
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

// The four fusion variants of one compiled shape.  L is a generic lambda
// taking the two flags as integral constants.
template <typename L>
hipError_t by_variant(const ImageFuse *fuse, const TaperFuse *taper, L &&launch)
{
    const bool tap = taper && taper->b > 0;
    if (fuse && tap) return launch(std::true_type{}, std::true_type{});
    if (fuse) return launch(std::true_type{}, std::false_type{});
    if (tap) return launch(std::false_type{}, std::true_type{});
    return launch(std::false_type{}, std::false_type{});
}

void fuse_args(const ImageFuse *fuse, const TaperFuse *taper, ImageFuse &fz, TaperFuse &tf)
{
    fz = ImageFuse{};
    tf = TaperFuse{};
    if (fuse) {
        fz = *fuse;
        if (fz.rdx <= 0) fz.rdx = 1;
        if (fz.rdy <= 0) fz.rdy = 1;
    }
    if (taper) tf = *taper;
}

template <typename T, int R, int TZ, int BY, int RY>
hipError_t launch_variant(const T *P, T *PP, const float *V, const Coefs &cf, const Grid &g,
                          const ImageFuse *fuse, const TaperFuse *taper, int *flag,
                          hipStream_t s, int xchunk)
{
    ImageFuse fz;
    TaperFuse tf;
    fuse_args(fuse, taper, fz, tf);
    return by_variant(fuse, taper, [&](auto img, auto tap) {
        return launch_cfg<T, R, TZ, BY, RY, decltype(img)::value, decltype(tap)::value>(
            P, PP, V, cf, g, fz, tf, flag, s, xchunk);
    });
}

template <typename T, int R, int BY, int RY, int NWM, bool IMG, bool TAPER>
hipError_t launch_rows_cfg(const T *P, T *PP, const float *V, const Coefs &cf, const Grid &g,
                           const ImageFuse &fz, const TaperFuse &tf, int *flag, hipStream_t s,
                           int xchunk)
{
    constexpr int TY = BY * RY;
    auto kern = k_stencil_rows<T, R, BY, RY, NWM, IMG, TAPER>;
    const int nw = (g.nz + 63) / 64;
    if (nw > NWM) return hipErrorInvalidConfiguration;
    RowGrid rg;
    rg.nty = (g.ny - 2 * R + TY - 1) / TY;
    const int nxi = g.nx - 2 * R;
    if (rg.nty <= 0 || nxi <= 0 || g.nz - 2 * R <= 0) return hipSuccess;
    const size_t lds = rows_lds_bytes<R, BY, RY, NWM>() + (TAPER ? (size_t)tf.b * sizeof(float) : 0);
    const int threads = 64 * nw * BY;
    const int cap = device_capacity((const void *)kern, threads, lds);
    if (cap <= 0) return hipErrorStreamCaptureUnsupported;
    if (xchunk <= 0) {
        const int n = choose_chunks(rg.nty, nxi, R, cap);
        xchunk = (nxi + n - 1) / n;
    }
    rg.xchunk = xchunk;
    rg.nxc = (nxi + xchunk - 1) / xchunk;
    hipLaunchKernelGGL(kern, dim3((unsigned)(rg.nty * rg.nxc)), dim3(64 * nw, BY), lds, s, P, PP,
                       V, cf, g, rg, fz, tf, tf.t, flag);
    return hipGetLastError();
}

template <typename T, int R, int BY, int RY, int NWM>
hipError_t launch_rows(const T *P, T *PP, const float *V, const Coefs &cf, const Grid &g,
                       const ImageFuse *fuse, const TaperFuse *taper, int *flag, hipStream_t s,
                       int xchunk)
{
    ImageFuse fz;
    TaperFuse tf;
    fuse_args(fuse, taper, fz, tf);
    return by_variant(fuse, taper, [&](auto img, auto tap) {
        return launch_rows_cfg<T, R, BY, RY, NWM, decltype(img)::value, decltype(tap)::value>(
            P, PP, V, cf, g, fz, tf, flag, s, xchunk);
    });
}

template <typename T, int R, int BY, int RY, int NWM, bool IMG, bool TAPER>
hipError_t launch_rows2_cfg(const T *P, T *PP, const float *V, const Coefs &cf, const Grid &g,
                            const ImageFuse &fz, const TaperFuse &tf, int *flag, hipStream_t s,
                            int xchunk)
{
    constexpr int TY = BY * RY;
    auto kern = k_stencil_rows2<T, R, BY, RY, NWM, IMG, TAPER>;
    const int nw = (g.nz + 127) / 128;
    if (nw > NWM || (g.nz & 1)) return hipErrorInvalidConfiguration;
    RowGrid rg;
    rg.nty = (g.ny - 2 * R + TY - 1) / TY;
    const int nxi = g.nx - 2 * R;
    if (rg.nty <= 0 || nxi <= 0 || g.nz - 2 * R <= 0) return hipSuccess;
    const size_t lds = 2u * (size_t)(TY + 2 * R) * (size_t)(128 * NWM + 2 * R) * sizeof(float) +
                       (TAPER ? (size_t)tf.b * sizeof(float) : 0);
    const int threads = 64 * nw * BY;
    const int cap = device_capacity((const void *)kern, threads, lds);
    if (cap <= 0) return hipErrorStreamCaptureUnsupported;
    if (xchunk <= 0) {
        const int n = choose_chunks(rg.nty, nxi, R, cap);
        xchunk = (nxi + n - 1) / n;
    }
    rg.xchunk = xchunk;
    rg.nxc = (nxi + xchunk - 1) / xchunk;
    hipLaunchKernelGGL(kern, dim3((unsigned)(rg.nty * rg.nxc)), dim3(64 * nw, BY), lds, s, P, PP,
                       V, cf, g, rg, fz, tf, tf.t, flag);
    return hipGetLastError();
}

template <typename T, int R, int BY, int RY, int NWM>
hipError_t launch_rows2(const T *P, T *PP, const float *V, const Coefs &cf, const Grid &g,
                        const ImageFuse *fuse, const TaperFuse *taper, int *flag, hipStream_t s,
                        int xchunk)
{
    ImageFuse fz;
    TaperFuse tf;
    fuse_args(fuse, taper, fz, tf);
    return by_variant(fuse, taper, [&](auto img, auto tap) {
        return launch_rows2_cfg<T, R, BY, RY, NWM, decltype(img)::value, decltype(tap)::value>(
            P, PP, V, cf, g, fz, tf, flag, s, xchunk);
    });
}

// One launch over the z tiles [zt0, zt0 + nzt); x chunks of at least
// minchunk planes.
template <typename T, int R, int RY, int WPB, bool IMG, int TAPER>
hipError_t launch_wave_cfg(const T *P, T *PP, const float *V, const Coefs &cf, const Grid &g,
                           const ImageFuse &fz, const TaperFuse &tf, int *flag, hipStream_t s,
                           int xchunk, int zt0, int nzt, int minchunk = 1)
{
    auto kern = k_stencil_wave<T, R, RY, WPB, IMG, TAPER>;
    WaveGrid wg;
    wg.zt0 = zt0;
    wg.nzt = nzt;
    if (nzt <= 0) return hipSuccess;
    wg.nyt = (g.ny - 2 * R + RY - 1) / RY;
    const int nxi = g.nx - 2 * R;
    if (wg.nyt <= 0 || nxi <= 0 || g.nz - 2 * R <= 0) return hipSuccess;
    const size_t lds = TAPER ? (size_t)tf.b * sizeof(float) : 0;
    if (xchunk <= 0) {
        const int cap = device_capacity((const void *)kern, 64 * WPB, lds) * WPB;
        if (cap <= 0) return hipErrorStreamCaptureUnsupported;
        const int n = choose_chunks((long long)wg.nzt * wg.nyt, nxi, R, cap);
        xchunk = (nxi + n - 1) / n;
    }
    xchunk = std::max(xchunk, minchunk);
    wg.xchunk = xchunk;
    wg.nxc = (nxi + xchunk - 1) / xchunk;
    wg.waves = (long long)wg.nzt * wg.nyt * wg.nxc;
    const long long nb = (wg.waves + WPB - 1) / WPB;
    hipLaunchKernelGGL(kern, dim3((unsigned)nb), dim3(64, WPB), lds, s, P, PP, V, cf, g, wg, fz,
                       tf, tf.t, flag);
    return hipGetLastError();
}

// The step with taper3D of both fields fused (TAPER 2, no imaging), on one-z
// waves over every z tile.  Front-right cells read plane b of P and PP: the
// byte offset from plane x < b to b must fit the 32-bit lane offset, and the
// first x chunk must reach plane b so that the wave reading PP there (x < b)
// is the one that updates it later.
template <typename T, int R, int RY, int WPB>
hipError_t launch_wave_all(const T *P, T *PP, const float *V, const Coefs &cf, const Grid &g,
                           const ImageFuse *fuse, const TaperFuse *taper, int *flag,
                           hipStream_t s, int xchunk)
{
    if (fuse || !taper || taper->b < R || taper->b > 8192) return hipErrorInvalidValue;
    if (2 * taper->b >= g.nx || 2 * taper->b >= g.ny || 2 * taper->b >= g.nz)
        return hipErrorInvalidValue;
    if ((long long)taper->b * g.sx * (long long)sizeof(T) + (long long)g.sy * g.ny * sizeof(T) >=
        (1LL << 31))
        return hipErrorInvalidValue;
    ImageFuse fz;
    TaperFuse tf;
    fuse_args(nullptr, taper, fz, tf);
    return launch_wave_cfg<T, R, RY, WPB, false, 2>(P, PP, V, cf, g, fz, tf, flag, s, xchunk, 0,
                                                    (g.nz + 63) / 64, taper->b - R + 1);
}

template <typename T, int R, int RY, int WPB>
hipError_t launch_wave(const T *P, T *PP, const float *V, const Coefs &cf, const Grid &g,
                       const ImageFuse *fuse, const TaperFuse *taper, int *flag, hipStream_t s,
                       int xchunk)
{
    ImageFuse fz;
    TaperFuse tf;
    fuse_args(fuse, taper, fz, tf);
    const int nzt = (g.nz + 63) / 64;
    // The upper taper touches only z < b: the z tiles whose cells and z halo
    // all lie at z >= b (zt * 64 - R >= b) run the untapered variant, with
    // fewer registers, in a second launch over disjoint cells.  16-bit only
    // (salt grid, FP16: tapered step 0.540 -> 0.398 ms, imaging step 0.83
    // -> 0.69); for FP32 the second launch costs more than it saves.
    const bool split = taper && !std::is_same<T, float>::value;
    const int ntap = split ? std::min(nzt, (taper->b + R + 63) / 64) : nzt;
    hipError_t e = by_variant(fuse, taper, [&](auto img, auto tap) {
        return launch_wave_cfg<T, R, RY, WPB, decltype(img)::value, decltype(tap)::value>(
            P, PP, V, cf, g, fz, tf, flag, s, xchunk, 0, ntap);
    });
    if (e != hipSuccess || ntap >= nzt) return e;
    return by_variant(fuse, nullptr, [&](auto img, auto tap) {
        return launch_wave_cfg<T, R, RY, WPB, decltype(img)::value, decltype(tap)::value>(
            P, PP, V, cf, g, fz, tf, flag, s, xchunk, ntap, nzt - ntap);
    });
}

// Two-cells-per-lane waves, one launch over the z tiles [zt0, zt0 + nzt) of
// 128 cells, tile t starting at z = zb + 128 t (zb even).
template <typename T, int R, int RY, int WPB, bool IMG, int TAPER>
hipError_t launch_wave2_cfg(const T *P, T *PP, const float *V, const Coefs &cf, const Grid &g,
                            const ImageFuse &fz, const TaperFuse &tf, int *flag, hipStream_t s,
                            int xchunk, int zb, int zt0, int nzt)
{
    auto kern = k_stencil_wave2<T, R, RY, WPB, IMG, TAPER>;
    WaveGrid wg;
    wg.zb = zb;
    wg.zt0 = zt0;
    wg.nzt = nzt;
    if (nzt <= 0) return hipSuccess;
    wg.nyt = (g.ny - 2 * R + RY - 1) / RY;
    const int nxi = g.nx - 2 * R;
    if (wg.nyt <= 0 || nxi <= 0 || g.nz - 2 * R <= 0) return hipSuccess;
    const size_t lds = TAPER ? (size_t)tf.b * sizeof(float) : 0;
    if (xchunk <= 0) {
        const int cap = device_capacity((const void *)kern, 64 * WPB, lds) * WPB;
        if (cap <= 0) return hipErrorStreamCaptureUnsupported;
        const int n = choose_chunks((long long)wg.nzt * wg.nyt, nxi, R, cap);
        xchunk = (nxi + n - 1) / n;
    }
    wg.xchunk = xchunk;
    wg.nxc = (nxi + xchunk - 1) / xchunk;
    wg.waves = (long long)wg.nzt * wg.nyt * wg.nxc;
    const long long nb = (wg.waves + WPB - 1) / WPB;
    hipLaunchKernelGGL(kern, dim3((unsigned)nb), dim3(64, WPB), lds, s, P, PP, V, cf, g, wg, fz,
                       tf, tf.t, flag);
    return hipGetLastError();
}

// Pair accesses need even nz and strides and pair-aligned arrays.
template <typename T>
bool wave2_fits(const T *P, const T *PP, const float *V, const ImageFuse *fuse, const Grid &g)
{
    auto al = [](const void *p, size_t a) { return ((uintptr_t)p % a) == 0; };
    if ((g.nz & 1) || (g.sy & 1) || (g.sx & 1)) return false;
    if (!al(P, 2 * sizeof(T)) || !al(PP, 2 * sizeof(T)) || !al(V, 8)) return false;
    if (fuse && (!al(fuse->ps, 2 * sizeof(T)) || !al(fuse->img, 8))) return false;
    return true;
}

template <typename T, int R, int RY, int WPB>
hipError_t launch_wave2(const T *P, T *PP, const float *V, const Coefs &cf, const Grid &g,
                        const ImageFuse *fuse, const TaperFuse *taper, int *flag, hipStream_t s,
                        int xchunk)
{
    ImageFuse fz;
    TaperFuse tf;
    fuse_args(fuse, taper, fz, tf);
    const int nzt = (g.nz + 127) / 128;
    // tapered z tiles first, the rest untapered (16-bit, as launch_wave)
    const bool split = taper && !std::is_same<T, float>::value;
    const int ntap = split ? std::min(nzt, (taper->b + R + 127) / 128) : nzt;
    hipError_t e = by_variant(fuse, taper, [&](auto img, auto tap) {
        return launch_wave2_cfg<T, R, RY, WPB, decltype(img)::value, decltype(tap)::value>(
            P, PP, V, cf, g, fz, tf, flag, s, xchunk, 0, 0, ntap);
    });
    if (e != hipSuccess || ntap >= nzt) return e;
    return by_variant(fuse, nullptr, [&](auto img, auto tap) {
        return launch_wave2_cfg<T, R, RY, WPB, decltype(img)::value, decltype(tap)::value>(
            P, PP, V, cf, g, fz, tf, flag, s, xchunk, 0, ntap, nzt - ntap);
    });
}

// The step with taper3D applied to its output (TAPER 3, no imaging): two-z
// 128 x 1 x 4 waves where cell pairs fit (even nz and strides), else one-z
// 64 x 1 x RY waves, over every z tile.
template <typename T, int R, int RY, int WPB>
hipError_t launch_wave_out(const T *P, T *PP, const float *V, const Coefs &cf, const Grid &g,
                           const ImageFuse *fuse, const TaperFuse *taper, int *flag,
                           hipStream_t s, int xchunk)
{
    if (fuse || !taper || taper->b < R || taper->b > 8192) return hipErrorInvalidValue;
    if (2 * taper->b >= g.nx || 2 * taper->b >= g.ny || 2 * taper->b >= g.nz)
        return hipErrorInvalidValue;
    ImageFuse fz;
    TaperFuse tf;
    fuse_args(nullptr, taper, fz, tf);
    if (wave2_fits(P, PP, V, nullptr, g))
        return launch_wave2_cfg<T, R, 4, WPB, false, 3>(P, PP, V, cf, g, fz, tf, flag, s, xchunk,
                                                        0, 0, (g.nz + 127) / 128);
    return launch_wave_cfg<T, R, RY, WPB, false, 3>(P, PP, V, cf, g, fz, tf, flag, s, xchunk, 0,
                                                    (g.nz + 63) / 64);
}

// The step with the upper taper applied to its output (TAPER 4), with or
// without imaging: one-z 64 x 1 x ry waves when the class tiling asks for
// them (ry 6 or 8), else two-z 128 x 1 x 4 waves where cell pairs fit, else
// one-z waves (64 x 1 x 6 with imaging, 64 x 1 x 8 without).
template <typename T, int R>
hipError_t launch_wave_outu(const T *P, T *PP, const float *V, const Coefs &cf, const Grid &g,
                            const ImageFuse *fuse, const TaperFuse *taper, int *flag,
                            hipStream_t s, const StencilTiling *t)
{
    if (!taper || taper->b < 1 || taper->b > 8192 || taper->hold < 0 || taper->hold > taper->b)
        return hipErrorInvalidValue;
    ImageFuse fz;
    TaperFuse tf;
    fuse_args(fuse, taper, fz, tf);
    const int xchunk = t ? t->xchunk : 0;
    const bool one = t && t->tz == 64 && t->by == 1 && (t->ry == 6 || t->ry == 8);
    const bool two = !one && wave2_fits(P, PP, V, fuse, g);
    const int ry = one ? t->ry : (fuse ? 6 : 8);
    if (fuse) {
        if (two)
            return launch_wave2_cfg<T, R, 4, 4, true, 4>(P, PP, V, cf, g, fz, tf, flag, s, xchunk,
                                                         0, 0, (g.nz + 127) / 128);
        if (ry == 8)
            return launch_wave_cfg<T, R, 8, 4, true, 4>(P, PP, V, cf, g, fz, tf, flag, s, xchunk,
                                                        0, (g.nz + 63) / 64);
        return launch_wave_cfg<T, R, 6, 4, true, 4>(P, PP, V, cf, g, fz, tf, flag, s, xchunk, 0,
                                                    (g.nz + 63) / 64);
    }
    if (two)
        return launch_wave2_cfg<T, R, 4, 4, false, 4>(P, PP, V, cf, g, fz, tf, flag, s, xchunk, 0,
                                                      0, (g.nz + 127) / 128);
    if (ry == 6)
        return launch_wave_cfg<T, R, 6, 4, false, 4>(P, PP, V, cf, g, fz, tf, flag, s, xchunk, 0,
                                                     (g.nz + 63) / 64);
    return launch_wave_cfg<T, R, 8, 4, false, 4>(P, PP, V, cf, g, fz, tf, flag, s, xchunk, 0,
                                                 (g.nz + 63) / 64);
}

// The 64-cell z tiles that the upper taper reaches (zt * 64 - R < b) on
// one-z waves of RY rows with the taper fused, every deeper cell on
// untapered two-z waves starting at the first untapered tile (even z).
// Without a taper: two-z waves throughout (the 16-bit default for untapered
// steps; salt grid, FP16: imaging 0.605 ms on one-z 6-row waves, 0.545 on
// two-z waves; plain 0.314 / 0.301).
template <typename T, int R, int RY>
hipError_t launch_wave_deep2(const T *P, T *PP, const float *V, const Coefs &cf, const Grid &g,
                             const ImageFuse *fuse, const TaperFuse *taper, int *flag,
                             hipStream_t s, int xchunk)
{
    ImageFuse fz;
    TaperFuse tf;
    fuse_args(fuse, taper, fz, tf);
    const bool tapered = taper && taper->b > 0;
    const int nzt = (g.nz + 63) / 64;
    const int ntap = tapered ? std::min(nzt, (taper->b + R + 63) / 64) : 0;
    if (ntap > 0) {
        const hipError_t e = by_variant(fuse, taper, [&](auto img, auto tap) {
            return launch_wave_cfg<T, R, RY, 4, decltype(img)::value, decltype(tap)::value>(
                P, PP, V, cf, g, fz, tf, flag, s, xchunk, 0, ntap);
        });
        if (e != hipSuccess || ntap >= nzt) return e;
    }
    const int zb = 64 * ntap;
    const int n2 = (g.nz - zb + 127) / 128;
    return by_variant(fuse, nullptr, [&](auto img, auto tap) {
        return launch_wave2_cfg<T, R, 4, 4, decltype(img)::value, decltype(tap)::value>(
            P, PP, V, cf, g, fz, tf, flag, s, xchunk, zb, 0, n2);
    });
}

// Full-row shapes compiled for radius 4 (by thread rows x ry rows each):
// (4, 4) up to 4 waves per row (nz <= 256); (2, 6) and (2, 7) up to 6 waves
// per row (nz <= 384).
inline bool rows_fit(int nz, int by, int ry)
{
    const int nw = (nz + 63) / 64;
    if (by == 4 && ry == 4) return nw <= 4;
    if (by == 2 && (ry == 6 || ry == 7)) return nw <= 6;
    return false;
}

// FP32 wavefields: every compiled shape.  16-bit wavefields: radius 4, the
// barrier-free waves and the full rows 4x4 / 2x6 on request, else the
// defaults (barrier-free waves: 64x1x6 for the imaging step, 64x1x8
// otherwise).
template <typename T, int R>
hipError_t launch_radius(const T *P, T *PP, const float *V, const Coefs &cf, const Grid &g,
                         const ImageFuse *fuse, const TaperFuse *taper, int *flag, hipStream_t s,
                         const StencilTiling *t)
{
    constexpr bool full = std::is_same<T, float>::value;
    const int xchunk = t ? t->xchunk : 0;
    if (taper && taper->all) {
        // taper3D fused: 64 x 1 x 8 one-z waves whatever the class tiling
        if constexpr (R == 4) {
            if (taper->all == 3)
                return launch_wave_outu<T, R>(P, PP, V, cf, g, fuse, taper, flag, s, t);
            if (taper->all == 2)
                return launch_wave_out<T, R, 8, 4>(P, PP, V, cf, g, fuse, taper, flag, s, xchunk);
            return launch_wave_all<T, R, 8, 4>(P, PP, V, cf, g, fuse, taper, flag, s, xchunk);
        }
        return hipErrorNotSupported;
    }
    if constexpr (R == 4 && std::is_same<T, float>::value) {
        // full rows with two z per lane (8-byte accesses): 2 x 4 x 3
        if (t && t->tz == 2 && t->by == 4 && t->ry == 3 && (g.nz & 1) == 0 &&
            (g.nz + 127) / 128 <= 3 && (!taper || taper->b <= 4096))
            return launch_rows2<T, R, 4, 3, 3>(P, PP, V, cf, g, fuse, taper, flag, s, xchunk);
    }
    if constexpr (R == 4) {
        // barrier-free waves, two z cells per lane: 128 x 1 x 4 (every storage
        // type); 128 x 1 x ry (ry 6, 8): one-z 64 x 1 x ry waves on the z tiles
        // the upper taper reaches, two-z waves below.  Odd nz or strides run
        // one cell per lane (64 x 1 x 8, or 64 x 1 x ry).
        if (t && t->tz == 128 && t->by == 1 && (!taper || taper->b <= 8192)) {
            const bool fits = wave2_fits(P, PP, V, fuse, g);
            if (t->ry == 4) {
                if (fits)
                    return launch_wave2<T, R, 4, 4>(P, PP, V, cf, g, fuse, taper, flag, s, xchunk);
                return launch_wave<T, R, 8, 4>(P, PP, V, cf, g, fuse, taper, flag, s, xchunk);
            }
            if (t->ry == 6)
                return fits ? launch_wave_deep2<T, R, 6>(P, PP, V, cf, g, fuse, taper, flag, s,
                                                         xchunk)
                            : launch_wave<T, R, 6, 4>(P, PP, V, cf, g, fuse, taper, flag, s,
                                                      xchunk);
            if (t->ry == 8)
                return fits ? launch_wave_deep2<T, R, 8>(P, PP, V, cf, g, fuse, taper, flag, s,
                                                         xchunk)
                            : launch_wave<T, R, 8, 4>(P, PP, V, cf, g, fuse, taper, flag, s,
                                                      xchunk);
        }
        // barrier-free waves: 64 x 1 x ry (every storage type)
        if (t && t->tz == 64 && t->by == 1 && (!taper || taper->b <= 8192)) {
            if (t->ry == 6)
                return launch_wave<T, R, 6, 4>(P, PP, V, cf, g, fuse, taper, flag, s, xchunk);
            if (t->ry == 8)
                return launch_wave<T, R, 8, 4>(P, PP, V, cf, g, fuse, taper, flag, s, xchunk);
        }
    }
    if constexpr (R == 4 && !std::is_same<T, float>::value) {
        // 16-bit wavefields: full rows (one z per lane) on request
        if (t && t->tz == 1 && (!taper || taper->b <= 4096)) {
            if (t->by == 4 && t->ry == 4 && rows_fit(g.nz, 4, 4))
                return launch_rows<T, R, 4, 4, 4>(P, PP, V, cf, g, fuse, taper, flag, s, xchunk);
            if (t->by == 2 && t->ry == 6 && rows_fit(g.nz, 2, 6))
                return launch_rows<T, R, 2, 6, 6>(P, PP, V, cf, g, fuse, taper, flag, s, xchunk);
        }
        // 16-bit wavefields by default: barrier-free waves.  Untapered steps
        // (plain, RBC imaging) on two-z 128 x 1 x 4 waves where pairs fit;
        // tapered ones on one-z waves (64 x 1 x 6 with imaging, 64 x 1 x 8
        // otherwise) split into tapered and untapered tiles.  Salt grid,
        // FP16: plain 0.305 ms (one-z 0.317), untapered imaging 0.545
        // (0.605), tapered 0.398, tapered imaging 0.672; two-z waves behind
        // the tapered tiles (launch_wave_deep2) made the HBC shot slower,
        // 4.49 s against 4.40 (the 272 deep cells fill three 128-cell tiles);
        // full rows 2 x 6: 0.546 / - / 0.562 / 0.744; pairs: 0.517 / - /
        // 0.563 / 0.778.
        if (!taper || taper->b <= 8192) {
            const bool tapered = taper && taper->b > 0;
            if (!tapered && wave2_fits(P, PP, V, fuse, g))
                return launch_wave_deep2<T, R, 8>(P, PP, V, cf, g, fuse, taper, flag, s, xchunk);
            if (fuse) return launch_wave<T, R, 6, 4>(P, PP, V, cf, g, fuse, taper, flag, s, xchunk);
            return launch_wave<T, R, 8, 4>(P, PP, V, cf, g, fuse, taper, flag, s, xchunk);
        }
    }
    if constexpr (R == 4 && std::is_same<T, float>::value) {
        // FP32 without a requested shape: the plain step on two-z 128 x 1 x 4
        // waves (64 x 1 x 8 waves where pairs do not fit), the tapered step
        // on 64 x 1 x 8 waves (salt grid: 0.412 and 0.463 ms against 0.447
        // and 0.498 for full rows 2 x 6); the fused receiver step keeps the
        // full rows below (0.699 ms against 0.79 or more on every wave shape).
        if (!t && !fuse && (!taper || taper->b <= 8192)) {
            const bool tapered = taper && taper->b > 0;
            if (!tapered && wave2_fits(P, PP, V, fuse, g))
                return launch_wave2<T, R, 4, 4>(P, PP, V, cf, g, fuse, taper, flag, s, xchunk);
            return launch_wave<T, R, 8, 4>(P, PP, V, cf, g, fuse, taper, flag, s, xchunk);
        }
    }
    if constexpr (R == 4) {
        // full-row blocks by default whenever a compiled shape covers nz (and
        // the taper table fits in LDS beside the plane buffers)
        // (a two-z-per-lane request that does not fit nz lands here too)
        const bool want_rows =
            (!t || t->tz == 1 || t->tz == 2 || !full) && (!taper || taper->b <= 4096);
        if (want_rows) {
            int by = t ? t->by : 0, ry = t ? t->ry : 0;
            if (!t || !full) {
                const bool keep = t && t->tz == 1 && (by == 4 && ry == 4 ? rows_fit(g.nz, 4, 4)
                                                      : by == 2 && ry == 6 && rows_fit(g.nz, 2, 6));
                if (!keep) {
                    by = ry = 0;
                    if (rows_fit(g.nz, 4, 4)) by = 4, ry = 4;
                    else if (rows_fit(g.nz, 2, 6)) by = 2, ry = 6;
                }
            }
            if (by == 4 && ry == 4 && rows_fit(g.nz, 4, 4))
                return launch_rows<T, R, 4, 4, 4>(P, PP, V, cf, g, fuse, taper, flag, s, xchunk);
            if (by == 2 && ry == 6 && rows_fit(g.nz, 2, 6))
                return launch_rows<T, R, 2, 6, 6>(P, PP, V, cf, g, fuse, taper, flag, s, xchunk);
            if constexpr (full) {
                if (by == 2 && ry == 7 && rows_fit(g.nz, 2, 7))
                    return launch_rows<T, R, 2, 7, 6>(P, PP, V, cf, g, fuse, taper, flag, s,
                                                      xchunk);
                // a full-row shape that does not cover nz: barrier-free
                // 64 x 1 x 8 waves (faster than the tiles at every measured
                // nz for the plain and the imaging step)
                if (!taper || taper->b <= 8192)
                    return launch_wave<T, R, 8, 4>(P, PP, V, cf, g, fuse, taper, flag, s, xchunk);
            }
            // a full-row shape that does not cover nz (16-bit): the same waves
            if constexpr (!full) {
                if (!taper || taper->b <= 8192)
                    return launch_wave<T, R, 8, 4>(P, PP, V, cf, g, fuse, taper, flag, s, xchunk);
            }
        }
    }
    if constexpr (full) {
        // the default tiles also serve the full-row, pair and wave shapes
        // where those do not apply (orders below 8, large tapers)
        if (!t || t->tz == 1 || t->tz == 2 || (t->tz == 64 && t->by == 1) || t->tz == 128 ||
            (t->tz == 64 && t->by == 4 && t->ry == 4))
            return launch_variant<T, R, 64, 4, 4>(P, PP, V, cf, g, fuse, taper, flag, s, xchunk);
        if (t->tz == 64 && t->by == 4 && t->ry == 2)
            return launch_variant<T, R, 64, 4, 2>(P, PP, V, cf, g, fuse, taper, flag, s, xchunk);
        if (t->tz == 64 && t->by == 8 && t->ry == 2)
            return launch_variant<T, R, 64, 8, 2>(P, PP, V, cf, g, fuse, taper, flag, s, xchunk);
        if (t->tz == 32 && t->by == 8 && t->ry == 4)
            return launch_variant<T, R, 32, 8, 4>(P, PP, V, cf, g, fuse, taper, flag, s, xchunk);
        if (t->tz == 32 && t->by == 8 && t->ry == 2)
            return launch_variant<T, R, 32, 8, 2>(P, PP, V, cf, g, fuse, taper, flag, s, xchunk);
        return hipErrorInvalidValue;
    } else {
        return launch_variant<T, R, 64, 4, 4>(P, PP, V, cf, g, fuse, taper, flag, s, xchunk);
    }
}

constexpr int kBatchWPB = 4;

template <typename T, int RY, bool IMG, bool TAPER>
hipError_t launch_batch_cfg(const BatchShot *shots, int n, long long nblocks, int par, int pspar,
                            const Coefs &cf, const TaperFuse &tf, int *flag, hipStream_t s)
{
    constexpr int R = 4;
    auto kern = k_stencil_batch<T, R, RY, kBatchWPB, IMG, TAPER>;
    const size_t lds = TAPER ? (size_t)tf.b * sizeof(float) : 0;
    if (nblocks <= 0) return hipSuccess;
    if (nblocks > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(kern, dim3((unsigned)nblocks), dim3(64, kBatchWPB), lds, s, shots, n,
                       nblocks, par, pspar, cf, tf, flag);
    return hipGetLastError();
}

template <typename T, int R, int RY, int WPB, bool TAPER>
hipError_t launch_dual_cfg(const T *Ps, T *PPs, const T *Pr, T *PPr, const float *V,
                           const Coefs &cf, const Grid &g, const ImageFuse &fz,
                           const TaperFuse &tf, int *flag, hipStream_t s, int xchunk)
{
    auto kern = k_stencil_dual<T, R, RY, WPB, TAPER>;
    WaveGrid wg;
    wg.zt0 = 0;
    wg.nzt = (g.nz + 63) / 64;
    wg.nyt = (g.ny - 2 * R + RY - 1) / RY;
    const int nxi = g.nx - 2 * R;
    if (wg.nyt <= 0 || nxi <= 0 || g.nz - 2 * R <= 0) return hipSuccess;
    const size_t lds = TAPER ? (size_t)tf.b * sizeof(float) : 0;
    if (xchunk <= 0) {
        const int cap = device_capacity((const void *)kern, 64 * WPB, lds) * WPB;
        if (cap <= 0) return hipErrorStreamCaptureUnsupported;
        const int n = choose_chunks((long long)wg.nzt * wg.nyt, nxi, R, cap);
        xchunk = (nxi + n - 1) / n;
    }
    wg.xchunk = xchunk;
    wg.nxc = (nxi + xchunk - 1) / xchunk;
    wg.waves = (long long)wg.nzt * wg.nyt * wg.nxc;
    const long long nb = (wg.waves + WPB - 1) / WPB;
    hipLaunchKernelGGL(kern, dim3((unsigned)nb), dim3(64, WPB), lds, s, Ps, PPs, Pr, PPr, V, cf,
                       g, wg, fz, tf, tf.t, flag);
    return hipGetLastError();
}

}  // namespace

bool dual_step_supported(int radius, int wave_format)
{
    return radius == 4 && wave_format == 0;
}

hipError_t launch_stencil_dual(int radius, const float *Ps, float *PPs, const float *Pr,
                               float *PPr, const float *V, const Coefs &cf, const Grid &g,
                               const ImageFuse &fuse, const TaperFuse *taper, int *flag,
                               hipStream_t s, const StencilTiling *tiling)
{
    if (radius != 4) return hipErrorNotSupported;
    const int xchunk = tiling ? tiling->xchunk : 0;
    ImageFuse fz;
    TaperFuse tf;
    fuse_args(&fuse, taper, fz, tf);
    if (taper && taper->b > 0)
        return launch_dual_cfg<float, 4, 4, 4, true>(Ps, PPs, Pr, PPr, V, cf, g, fz, tf, flag, s,
                                                     xchunk);
    return launch_dual_cfg<float, 4, 4, 4, false>(Ps, PPs, Pr, PPr, V, cf, g, fz, tf, flag, s,
                                                  xchunk);
}

bool stencil_tiling_supported(int tz, int by, int ry)
{
    return (tz == 64 && by == 4 && (ry == 4 || ry == 2)) || (tz == 64 && by == 8 && ry == 2) ||
           (tz == 32 && by == 8 && (ry == 4 || ry == 2)) ||
           (tz == 1 && ((by == 4 && ry == 4) || (by == 2 && (ry == 6 || ry == 7)))) ||
           (tz == 64 && by == 1 && (ry == 6 || ry == 8)) || (tz == 2 && by == 4 && ry == 3) ||
           (tz == 128 && by == 1 && (ry == 4 || ry == 6 || ry == 8));
}

template <typename T>
hipError_t launch_stencil(int radius, const T *P, T *PP, const float *V, const Coefs &cf,
                          const Grid &g, const ImageFuse *fuse, const TaperFuse *taper,
                          int *flag, hipStream_t s, const StencilTiling *tiling)
{
    if constexpr (std::is_same<T, float>::value) {
        switch (radius) {
        case 1: return launch_radius<T, 1>(P, PP, V, cf, g, fuse, taper, flag, s, tiling);
        case 2: return launch_radius<T, 2>(P, PP, V, cf, g, fuse, taper, flag, s, tiling);
        case 3: return launch_radius<T, 3>(P, PP, V, cf, g, fuse, taper, flag, s, tiling);
        case 4: return launch_radius<T, 4>(P, PP, V, cf, g, fuse, taper, flag, s, tiling);
        default: return hipErrorInvalidValue;
        }
    } else {
        if (radius != 4) return hipErrorNotSupported;
        return launch_radius<T, 4>(P, PP, V, cf, g, fuse, taper, flag, s, tiling);
    }
}

void capture_guard(bool on) { g_capture_depth += on ? 1 : -1; }

long long plan_stencil_batch(int ry, BatchShot *hs, int n, int R, int blen)
{
    if (R != 4 || (ry != 6 && ry != 8) || n <= 0) return -1;
    // x chunks: one count for the batch, from the total tile count and the
    // resident-wave capacity of the kernel
    const void *fn = ry == 6 ? (const void *)k_stencil_batch<float, 4, 6, kBatchWPB, true, true>
                             : (const void *)k_stencil_batch<float, 4, 8, kBatchWPB, true, true>;
    const int cap = device_capacity(fn, 64 * kBatchWPB, (size_t)blen * sizeof(float)) * kBatchWPB;
    if (cap <= 0) return -1;
    long long tiles = 0;
    int nxmin = 1 << 30;
    for (int i = 0; i < n; ++i) {
        const Grid &g = hs[i].g;
        hs[i].nzt = (g.nz + 63) / 64;
        hs[i].nyt = (g.ny - 2 * R + ry - 1) / ry;
        if (hs[i].nyt <= 0 || g.nx - 2 * R <= 0 || g.nz - 2 * R <= 0) return -1;
        tiles += (long long)hs[i].nzt * hs[i].nyt;
        nxmin = std::min(nxmin, g.nx - 2 * R);
    }
    const int nc = choose_chunks(tiles, nxmin, R, cap);
    long long blk = 0;
    for (int i = 0; i < n; ++i) {
        const int nxi = hs[i].g.nx - 2 * R;
        hs[i].xchunk = (nxi + nc - 1) / nc;
        const long long nxc = (nxi + hs[i].xchunk - 1) / hs[i].xchunk;
        hs[i].nwaves = (long long)hs[i].nzt * hs[i].nyt * nxc;
        hs[i].blk0 = blk;
        blk += (hs[i].nwaves + kBatchWPB - 1) / kBatchWPB;
    }
    return blk;
}

template <typename T>
hipError_t launch_stencil_batch(int ry, const BatchShot *shots, const BatchShot *hs, int n,
                                long long nblocks, bool img, int par, int pspar, const Coefs &cf,
                                const TaperFuse *taper, int *flag, hipStream_t s)
{
    (void)hs;
    TaperFuse tf{};
    if (taper) tf = *taper;
    else tf.b = 0;
    if (img) tf.b = taper ? taper->b : tf.b;
    auto go = [&](auto ryc) -> hipError_t {
        constexpr int RY = decltype(ryc)::value;
        if (img && taper)
            return launch_batch_cfg<T, RY, true, true>(shots, n, nblocks, par, pspar, cf, tf, flag,
                                                       s);
        if (taper)
            return launch_batch_cfg<T, RY, false, true>(shots, n, nblocks, par, pspar, cf, tf,
                                                        flag, s);
        if (img) return hipErrorInvalidValue;  // the batch images HBC steps only
        return launch_batch_cfg<T, RY, false, false>(shots, n, nblocks, par, pspar, cf, tf, flag,
                                                     s);
    };
    if (ry == 6) return go(std::integral_constant<int, 6>{});
    if (ry == 8) return go(std::integral_constant<int, 8>{});
    return hipErrorInvalidValue;
}

#define RTM_INSTANTIATE_BATCH(T)                                                                 \
    template hipError_t launch_stencil_batch<T>(int, const BatchShot *, const BatchShot *, int,  \
                                                long long, bool, int, int, const Coefs &,        \
                                                const TaperFuse *, int *, hipStream_t);
RTM_INSTANTIATE_BATCH(float)
RTM_INSTANTIATE_BATCH(_Float16)
RTM_INSTANTIATE_BATCH(__bf16)

template hipError_t launch_stencil<float>(int, const float *, float *, const float *,
                                          const Coefs &, const Grid &, const ImageFuse *,
                                          const TaperFuse *, int *, hipStream_t,
                                          const StencilTiling *);
template hipError_t launch_stencil<_Float16>(int, const _Float16 *, _Float16 *, const float *,
                                             const Coefs &, const Grid &, const ImageFuse *,
                                             const TaperFuse *, int *, hipStream_t,
                                             const StencilTiling *);
template hipError_t launch_stencil<__bf16>(int, const __bf16 *, __bf16 *, const float *,
                                           const Coefs &, const Grid &, const ImageFuse *,
                                           const TaperFuse *, int *, hipStream_t,
                                           const StencilTiling *);

}  // namespace rtmhip
