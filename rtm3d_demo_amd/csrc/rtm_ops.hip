// rtm_ops.hip -- border, source, receiver, upper-border and imaging kernels
// for gfx950.  Each restates one operator of the reference CPU platform
// (rtm-3d/src/platform/host/RTMCPUPlatform.cpp) with the same per-cell
// arithmetic; they are bandwidth-trivial next to the stencil and are kept
// simple and coalesced (z on consecutive lanes).  Wavefields have storage
// type T (float, _Float16, __bf16): values are widened on load, computed in
// float and rounded once per stored value, as the reference's in-place
// operators would with grids of that type.
#include "rtm_internal.hpp"

#include <algorithm>

#pragma clang fp contract(off)

namespace rtmhip {
namespace {

constexpr int kThreads = 256;

// Division of a dividend 0 <= n < 2^31 by a divisor d >= 1 fixed for a launch:
// q = umulhi(n, m) >> s with l = ceil(log2 d), m = ceil(2^(31+l) / d) and
// s = l - 1 (Granlund and Montgomery, "Division by invariant integers using
// multiplication", PLDI 1994, with N = 31: m < 2^32 and the result is exact
// for every such n); d = 1 is passed through.  A 32-bit division costs tens of
// instructions per lane, which made the border kernels ALU-bound.
struct FastDiv {
    uint32_t m;
    int s;  // -1 for d == 1
};

inline FastDiv fast_div(uint32_t d)
{
    if (d <= 1) return FastDiv{0u, -1};
    int l = 0;
    while ((1u << l) < d) ++l;
    const uint64_t m = ((uint64_t(1) << (31 + l)) + d - 1) / d;
    return FastDiv{(uint32_t)m, l - 1};
}

__device__ __forceinline__ int fdiv(int n, FastDiv d)
{
    return d.s < 0 ? n : (int)(__umulhi((uint32_t)n, d.m) >> d.s);
}

__device__ __forceinline__ long long goff(const Grid &g, int x, int y, int z)
{
    return (long long)x * g.sx + (long long)y * g.sy + z;
}

// Taper factor along one axis of length n: g[b-1-i] in the low border,
// g[i-(n-b)] in the high border.
__device__ __forceinline__ float gfac(const float *__restrict__ t, int i, int n, int b)
{
    return i < b ? t[b - 1 - i] : t[i - (n - b)];
}

// taperUpper3D (RTMCPUPlatform.cpp:566-657): every cell with z < blen, of
// one or two grids (blockIdx.y picks the grid), as ((v * f1) * f2) * f3:
//   x, y interior           : v*gz
//   x in a border (any y)   : ((v*gy if y in a border)*gx)*gz
//   y in a border, x inner  : (v*gy)*gz
// Cells are numbered z fastest over [0, nx*ny*b); every thread handles
// kUpperPer of them kThreads apart and issues their loads first.
constexpr int kUpperPer = 4;

template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_taper_upper(T *__restrict__ P0, T *__restrict__ P1, const float *__restrict__ t, Grid g,
                  int b, int n, FastDiv db, FastDiv dny)
{
    T *__restrict__ P = blockIdx.y ? P1 : P0;
    const int i0 = (int)blockIdx.x * (kThreads * kUpperPer) + (int)threadIdx.x;
    long long off[kUpperPer];
    float f1[kUpperPer], f2[kUpperPer], f3[kUpperPer], v[kUpperPer];
#pragma unroll
    for (int k = 0; k < kUpperPer; ++k) {
        const int i = i0 + k * kThreads;
        off[k] = -1;
        v[k] = 0.0f;
        if (i >= n) continue;
        const int r = fdiv(i, db);
        const int z = i - r * b;
        const int x = fdiv(r, dny);
        const int y = r - x * g.ny;
        const bool bx = x < b || x >= g.nx - b;
        const bool by = y < b || y >= g.ny - b;
        const float gy = by ? gfac(t, y, g.ny, b) : 1.0f;
        const float gz = t[b - 1 - z];
        f1[k] = bx ? gy : (by ? gy : gz);
        f2[k] = bx ? gfac(t, x, g.nx, b) : (by ? gz : 1.0f);
        f3[k] = bx ? gz : 1.0f;
        off[k] = (long long)x * g.sx + (long long)y * g.sy + z;
        v[k] = wld(P[off[k]]);
    }
#pragma unroll
    for (int k = 0; k < kUpperPer; ++k) {
        if (off[k] < 0) continue;
        float w = v[k] * f1[k];
        w *= f2[k];
        w *= f3[k];
        P[off[k]] = wst<T>(w);
    }
}

// 16-byte groups of G = 16/sizeof(T) consecutive z cells, for the border
// kernels' vector forms: one 16-byte load and store per lane.  Used when the
// border width and nz are multiples of G and the fields are 16-byte aligned
// (then no group straddles a border edge).
template <typename T>
struct alignas(16) Group16 {
    static constexpr int G = 16 / (int)sizeof(T);
    T v[G];
};

template <typename T>
__device__ __forceinline__ void taper_group(T *__restrict__ p, const float *__restrict__ t,
                                            int z0, int nz, int b, float f1, float f2,
                                            bool bz_factor, int zmode)
{
    // zmode: 0 no z factor; 1 z factor after f1 and f2 ((v*f1)*f2)*gz;
    //        2 z factor in second place (v*f1)*gz
    Group16<T> a = *reinterpret_cast<Group16<T> *>(p);
#pragma unroll
    for (int k = 0; k < Group16<T>::G; ++k) {
        const int z = z0 + k;
        const float gz = bz_factor ? gfac(t, z, nz, b) : 1.0f;
        float w = wld(a.v[k]) * f1;
        if (zmode == 2) {
            w *= gz;
        } else {
            w *= f2;
            if (zmode == 1) w *= gz;
        }
        a.v[k] = wst<T>(w);
    }
    *reinterpret_cast<Group16<T> *>(p) = a;
}

// taperUpper3D over 16-byte groups: the band z < b of every (x, y) row.
template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_taper_upper_v(T *__restrict__ P0, T *__restrict__ P1, const float *__restrict__ t, Grid g,
                    int b, int ngroups, FastDiv dbg, FastDiv dny)
{
    constexpr int G = Group16<T>::G;
    T *__restrict__ P = blockIdx.y ? P1 : P0;
    const int q = (int)(blockIdx.x * kThreads + threadIdx.x);
    if (q >= ngroups) return;
    const int bg = b / G;
    const int r = fdiv(q, dbg);
    const int z0 = (q - r * bg) * G;
    const int x = fdiv(r, dny);
    const int y = r - x * g.ny;
    const bool bx = x < b || x >= g.nx - b;
    const bool by = y < b || y >= g.ny - b;
    const float gy = by ? gfac(t, y, g.ny, b) : 1.0f;
    // x border: ((v*gy)*gx)*gz; y border: (v*gy)*gz; inner: v*gz
    const float f1 = bx || by ? gy : 1.0f;
    const float f2 = bx ? gfac(t, x, g.nx, b) : 1.0f;
    taper_group(P + ((long long)x * g.sx + (long long)y * g.sy + z0), t, z0, g.nz, b, f1, f2,
                true, 1);
}

// taper3D over 16-byte groups (the front-right edge's interior-z cells
// excepted): the z-border groups of the rows with x and y interior, then
// every group of the rows with x or y in a border, as k_taper_all.
template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_taper_all_v(T *__restrict__ P0, T *__restrict__ P1, const float *__restrict__ t, Grid g,
                  int b, int nA, int nB, FastDiv d2bg, FastDiv dly, FastDiv dnzg, FastDiv dny,
                  FastDiv d2b)
{
    constexpr int G = Group16<T>::G;
    T *__restrict__ P = blockIdx.y ? P1 : P0;
    const int i = (int)(blockIdx.x * kThreads + threadIdx.x);
    const int b2 = 2 * b, bg2 = b2 / G, nzg = g.nz / G;
    int x, y, z0;
    bool bx, by;
    if (i < nA) {
        const int r = fdiv(i, d2bg), j = (i - r * bg2) * G;
        const int xi = fdiv(r, dly);
        x = b + xi;
        y = b + (r - xi * (g.ny - b2));
        z0 = j < b ? j : g.nz - b2 + j;
        bx = by = false;
    } else if (i - nA < nB) {
        const int e = i - nA;
        const int r = fdiv(e, dnzg);
        z0 = (e - r * nzg) * G;
        const int nxs = b2 * g.ny;
        if (r < nxs) {
            const int xx = fdiv(r, dny);
            y = r - xx * g.ny;
            x = xx < b ? xx : g.nx - b2 + xx;
            bx = true;
            by = y < b || y >= g.ny - b;
        } else {
            const int q = r - nxs, xi = fdiv(q, d2b), yy = q - xi * b2;
            x = b + xi;
            y = yy < b ? yy : g.ny - b2 + yy;
            bx = false;
            by = true;
        }
    } else {
        return;
    }
    const bool bz = z0 < b || z0 >= g.nz - b;  // the whole group (b, nz multiples of G)
    if (bx && by && !bz && x < b && y >= g.ny - b) return;  // front-right edge
    const float gx = bx ? gfac(t, x, g.nx, b) : 1.0f;
    const float gy = by ? gfac(t, y, g.ny, b) : 1.0f;
    T *p = P + ((long long)x * g.sx + (long long)y * g.sy + z0);
    if (bz) {
        // (bx: [gy] gx gz) | (by: gy gz) | gz
        if (bx) taper_group(p, t, z0, g.nz, b, gy, gx, true, 1);
        else if (by) taper_group(p, t, z0, g.nz, b, gy, 1.0f, true, 2);
        else taper_group(p, t, z0, g.nz, b, 1.0f, 1.0f, true, 2);
    } else {
        // (bx and by: gx gy) | gy | gx
        taper_group(p, t, z0, g.nz, b, bx ? gx : gy, bx && by ? gy : 1.0f, false, 0);
    }
}

// The slab cells (z < b) that the stencil does not update: z < R, or x / y
// within R of the grid edge.  frame_cell decodes the i-th of them (z < R
// first, then the x frame, then the y frame); false past the last.
__device__ __forceinline__ long long frame_count(const Grid &g, int b, int R)
{
    const long long lz = b - R;
    return (long long)g.nx * g.ny * R + (long long)2 * R * g.ny * lz +
           (long long)(g.nx - 2 * R) * 2 * R * lz;
}
__device__ __forceinline__ bool in_frame(const Grid &g, int b, int R, int x, int y, int z)
{
    return z < b && (z < R || x < R || x >= g.nx - R || y < R || y >= g.ny - R);
}
template <typename I>
__device__ __forceinline__ bool frame_cell(const Grid &g, int b, int R, I i, int &x, int &y,
                                           int &z)
{
    const I nA = (I)g.nx * g.ny * R;  // z < R, all x, y
    const int lz = b - R;             // z in [R, b)
    const I nB = (I)2 * R * g.ny * lz;              // x frame
    const I nC = (I)(g.nx - 2 * R) * 2 * R * lz;    // y frame
    if (i < nA) {
        z = (int)(i % R);
        const I r = i / R;
        y = (int)(r % g.ny);
        x = (int)(r / g.ny);
    } else if ((i -= nA) < nB) {
        z = R + (int)(i % lz);
        const I r = i / lz;
        y = (int)(r % g.ny);
        const int xx = (int)(r / g.ny);
        x = xx < R ? xx : g.nx - 2 * R + xx;
    } else if ((i -= nB) < nC) {
        z = R + (int)(i % lz);
        const I r = i / lz;
        const int yy = (int)(r % (2 * R));
        x = R + (int)(r / (2 * R));
        y = yy < R ? yy : g.ny - 2 * R + yy;
    } else {
        return false;
    }
    return true;
}
// The upper taper of one slab cell applied `taps` times in place: each
// application is one in-place pass of the reference, rounded to T before
// the next; the last rounding is the caller's store.
template <typename T>
__device__ __forceinline__ float frame_taper(float v, const float *__restrict__ t, const Grid &g,
                                             int b, int x, int y, int z, int taps)
{
    const bool bx = x < b || x >= g.nx - b;
    const bool by = y < b || y >= g.ny - b;
    const float fy = by ? gfac(t, y, g.ny, b) : 1.0f;
    const float fx = bx ? gfac(t, x, g.nx, b) : 1.0f;
    const float fz = t[b - 1 - z];
    for (int k = 0; k < taps; ++k) {
        if (k) v = wrnd<T>(v);
        v *= fy;
        v *= fx;
        v *= fz;
    }
    return v;
}

// frame_cell and image_frame_cell with 32-bit indices and divisions by launch constants
// through a multiply-high (the single-grid kernels: a 64-bit division per
// cell made the receiver-and-frame imaging kernel ALU-bound).  dv: divisors
// {R, z extent of the frame, ny, 2R}.
struct FrameDivs {
    FastDiv r, lz, ny, r2;
};
__device__ __forceinline__ bool frame_cell_fd(const Grid &g, int b, int R, int i,
                                              const FrameDivs &dv, int &x, int &y, int &z)
{
    const int nA = g.nx * g.ny * R;
    const int lz = b - R;
    const int nB = 2 * R * g.ny * lz;
    const int nC = (g.nx - 2 * R) * 2 * R * lz;
    if (i < nA) {
        const int r = fdiv(i, dv.r);
        z = i - r * R;
        x = fdiv(r, dv.ny);
        y = r - x * g.ny;
    } else if ((i -= nA) < nB) {
        const int r = fdiv(i, dv.lz);
        z = R + (i - r * lz);
        const int xx = fdiv(r, dv.ny);
        y = r - xx * g.ny;
        x = xx < R ? xx : g.nx - 2 * R + xx;
    } else if ((i -= nB) < nC) {
        const int r = fdiv(i, dv.lz);
        z = R + (i - r * lz);
        const int xr = fdiv(r, dv.r2);
        const int yy = r - xr * 2 * R;
        x = R + xr;
        y = yy < R ? yy : g.ny - 2 * R + yy;
    } else {
        return false;
    }
    return true;
}
__device__ __forceinline__ bool image_frame_cell_fd(const Grid &g, int R, int z0, int z1, int i,
                                                    const FrameDivs &dv, int &x, int &y, int &z)
{
    const int lz = z1 - z0;
    if (lz <= 0) return false;
    const int nA = 2 * R * g.ny * lz;
    const int nB = (g.nx - 2 * R) * 2 * R * lz;
    if (i < nA) {
        const int r = fdiv(i, dv.lz);
        z = z0 + (i - r * lz);
        const int xx = fdiv(r, dv.ny);
        y = r - xx * g.ny;
        x = xx < R ? xx : g.nx - 2 * R + xx;
    } else if ((i -= nA) < nB) {
        const int r = fdiv(i, dv.lz);
        z = z0 + (i - r * lz);
        const int xr = fdiv(r, dv.r2);
        const int yy = r - xr * 2 * R;
        x = R + xr;
        y = yy < R ? yy : g.ny - 2 * R + yy;
    } else {
        return false;
    }
    return true;
}
inline FrameDivs frame_divs(int R, int lz, int ny)
{
    return FrameDivs{fast_div((uint32_t)R), fast_div((uint32_t)std::max(lz, 1)),
                     fast_div((uint32_t)ny), fast_div((uint32_t)(2 * R))};
}

// In-place upper taper, applied `taps` times, of the frame cells.  The fused
// HBC step tapers every other slab value as it is loaded; these cells of PP
// are not read during the step, so tapering them in place here keeps PP's
// memory equal to the reference's after its in-place tapers.
template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_taper_frame(T *__restrict__ P, const float *__restrict__ t, Grid g, int b, int R, int taps,
                  FrameDivs dv, int zhold)
{
    const int i = (int)(blockIdx.x * kThreads + threadIdx.x);  // count < 2^31 (launcher)
    int x, y, z;
    if (!frame_cell_fd(g, b, R, i, dv, x, y, z)) return;
    const long long o = goff(g, x, y, z);
    P[o] = wst<T>(frame_taper<T>(wld(P[o]), t, g, b, x, y, z, z >= zhold ? taps - 1 : taps));
}

// taper3D at one cell, `taps` times (rounded to T between applications):
// on the z border ((v*fy)*fx)*fz, off it (v*fx)*fy -- the x-y edge pass's
// order, and the face passes' with the other factor 1 (exact).
template <typename T>
__device__ __forceinline__ float all_taper(float v, const float *__restrict__ t, const Grid &g,
                                           int b, int x, int y, int z, int taps)
{
    const bool bx = x < b || x >= g.nx - b;
    const bool by = y < b || y >= g.ny - b;
    const bool bz = z < b || z >= g.nz - b;
    const float fx = bx ? gfac(t, x, g.nx, b) : 1.0f;
    const float fy = by ? gfac(t, y, g.ny, b) : 1.0f;
    const float fz = bz ? gfac(t, z, g.nz, b) : 1.0f;
    for (int k = 0; k < taps; ++k) {
        if (k) v = wrnd<T>(v);
        if (bz) {
            v *= fy;
            v *= fx;
            v *= fz;
        } else {
            v *= fx;
            v *= fy;
        }
    }
    return v;
}

// The frame of the fused every-border step: the cells with x or y within R
// of a grid face, every z (the z frame of the other rows is stored by the
// stencil itself), except the front-right edge cells (x < b, y >= ny-b, z
// inside), tapered `taps` times in place.  Index: the x frame (2R planes),
// then the y frame of the inner planes; z fastest.
template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_taper_frame_all(T *__restrict__ P, const float *__restrict__ t, Grid g, int b, int R,
                      int taps, int nX, int n)
{
    int i = (int)(blockIdx.x * kThreads + threadIdx.x);  // n < 2^31 (launcher)
    if (i >= n) return;
    const int z = i % g.nz;
    int x, y;
    if (i < nX) {
        const int r = i / g.nz;
        y = r % g.ny;
        const int xx = r / g.ny;
        x = xx < R ? xx : g.nx - 2 * R + xx;
    } else {
        const int r = (i - nX) / g.nz;
        const int yy = r % (2 * R);
        x = R + r / (2 * R);
        y = yy < R ? yy : g.ny - 2 * R + yy;
    }
    if (x < b && y >= g.ny - b && z >= b && z < g.nz - b) return;  // k_taper_frame_quirk
    const long long o = goff(g, x, y, z);
    P[o] = wst<T>(all_taper<T>(wld(P[o]), t, g, b, x, y, z, taps));
}

// The front-right edge of taper3D applied to a step's output (the step
// fused with TAPER 3 stored every other cell tapered): cells x < b, y >=
// ny-b, z in [b, nz-b) take the face-tapered value of (b, y, z), which the
// step stored as (v*gy) rounded, then ((.)*gx)*gy.  Runs after the step and
// after the frame pass.
template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_taper_quirk_out(T *__restrict__ P, const float *__restrict__ t, Grid g, int b, FastDiv dlz,
                      FastDiv db, int n)
{
    const int i = (int)(blockIdx.x * kThreads + threadIdx.x);  // n < 2^31 (launcher)
    if (i >= n) return;
    const int lz = g.nz - 2 * b;
    const int r = fdiv(i, dlz);
    const int z = b + (i - r * lz);
    const int x = fdiv(r, db);
    const int iy = r - x * b;
    const int y = g.ny - b + iy;
    const float gy = t[iy];
    float v = wld(P[goff(g, b, y, z)]);
    v *= t[b - 1 - x];
    v *= gy;
    P[goff(g, x, y, z)] = wst<T>(v);
}

// The front-right edge cells of that frame (x < R, or y >= ny-R), from the
// cell (b, y, z) before this step's tapers: the face pass's (v*gy) rounded,
// `taps` times, then (v*gx)*gy.  Runs before k_taper_frame_all, which
// rescales the frame cells of plane b.
template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_taper_frame_quirk(T *__restrict__ P, const float *__restrict__ t, Grid g, int b, int R,
                        int taps, int n1, int n)
{
    const int i = (int)(blockIdx.x * kThreads + threadIdx.x);  // n < 2^31 (launcher)
    if (i >= n) return;
    const int lz = g.nz - 2 * b;
    const int z = b + i % lz;
    int x, iy;
    if (i < n1) {  // x < R, y in [ny-b, ny)
        const int r = i / lz;
        iy = r % b;
        x = r / b;
    } else {  // x in [R, b), y in [ny-R, ny)
        const int r = (i - n1) / lz;
        iy = b - R + r % R;
        x = R + r / R;
    }
    const int y = g.ny - b + iy;
    const float gy = t[iy];
    float v = wld(P[goff(g, b, y, z)]);
    for (int k = 0; k < taps; ++k) v = wrnd<T>(v * gy);
    v *= t[b - 1 - x];
    v *= gy;
    P[goff(g, x, y, z)] = wst<T>(v);
}

// The front-right edge of taper3D (RTMCPUPlatform.cpp:900-904) writes cells
// x < blen, y >= ny-blen, z interior from the face-tapered cell
// (blen, y, z).  This kernel runs before k_taper_all so it reads that cell
// before the face pass rescales it:  ((P(blen,y,z)*gy)*gx)*gy, where the
// face pass's product P*gy is the value it stored (rounded to T).
template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_taper_front_right(T *__restrict__ P0, T *__restrict__ P1, const float *__restrict__ t,
                        Grid g, int b)
{
    T *__restrict__ P = blockIdx.y ? P1 : P0;  // one grid per blockIdx.y
    const int lz = g.nz - 2 * b;
    const long long n = (long long)b * b * lz;
    const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const int z = b + (int)(i % lz);
    const long long r = i / lz;
    const int iy = (int)(r % b);
    const int x = (int)(r / b);
    const int y = g.ny - b + iy;
    const float gy = t[iy];
    float v = wld(P[goff(g, b, y, z)]);
    v *= gy;
    v = wrnd<T>(v);
    v *= t[b - 1 - x];
    v *= gy;
    P[goff(g, x, y, z)] = wst<T>(v);
}

// taper3D (RTMCPUPlatform.cpp:710-920) for every border cell except the
// front-right edge's interior-z cells, of one or two grids (blockIdx.y picks
// the grid), over two packed index ranges:
//   [0, nA)        the z-border cells (z < b or z >= nz-b) of the rows with x
//                  and y interior: 2b consecutive cells per row, z factor only;
//   [nA, nA + nB)  every cell of the rows with x or y in a border (the x
//                  border slabs for all y first, then the y border rows of
//                  the interior x), nz per row.
// Each cell gets up to three factors in the reference loop nests' order,
// ((v * f1) * f2) * f3, unused ones being 1 (v * 1 == v exactly).  Every
// thread handles kPer cells kThreads apart and issues their loads first.
constexpr int kTaperPer = 4;

struct TaperCell {
    long long off;
    float f1, f2, f3;
    bool ok;
};

struct TaperDivs {
    FastDiv b2, ly, nz, ny;  // by 2b, ny-2b, nz, ny
};

__device__ __forceinline__ TaperCell taper_cell(const float *__restrict__ t, const Grid &g, int b,
                                                int nA, int nB, const TaperDivs &dv, int i)
{
    TaperCell c{0, 1.0f, 1.0f, 1.0f, false};
    const int b2 = 2 * b;
    int x, y, z;
    bool bx, by;
    if (i < nA) {
        const int ly = g.ny - b2;
        const int r = fdiv(i, dv.b2), j = i - r * b2;
        const int xi = fdiv(r, dv.ly);
        x = b + xi;
        y = b + (r - xi * ly);
        z = j < b ? j : g.nz - b2 + j;
        bx = by = false;
    } else if (i - nA < nB) {
        const int e = i - nA;
        const int r = fdiv(e, dv.nz);
        z = e - r * g.nz;
        const int nxs = b2 * g.ny;  // rows of the two x border slabs
        if (r < nxs) {
            const int xx = fdiv(r, dv.ny);
            y = r - xx * g.ny;
            x = xx < b ? xx : g.nx - b2 + xx;
            bx = true;
            by = y < b || y >= g.ny - b;
        } else {
            const int q = r - nxs, xi = fdiv(q, dv.b2), yy = q - xi * b2;
            x = b + xi;
            y = yy < b ? yy : g.ny - b2 + yy;
            bx = false;
            by = true;
        }
    } else {
        return c;
    }
    const bool bz = z < b || z >= g.nz - b;
    if (bx && by && !bz && x < b && y >= g.ny - b) return c;  // front-right edge
    const float gx = bx ? gfac(t, x, g.nx, b) : 1.0f;
    const float gy = by ? gfac(t, y, g.ny, b) : 1.0f;
    const float gz = bz ? gfac(t, z, g.nz, b) : 1.0f;
    if (bz) {
        // (bx: [gy] gx gz) | (by: gy gz) | gz
        c.f1 = bx ? gy : (by ? gy : gz);
        c.f2 = bx ? gx : (by ? gz : 1.0f);
        c.f3 = bx ? gz : 1.0f;
    } else {
        // (bx and by: gx gy) | gy | gx
        c.f1 = bx ? gx : gy;
        c.f2 = bx && by ? gy : 1.0f;
    }
    c.off = (long long)x * g.sx + (long long)y * g.sy + z;
    c.ok = true;
    return c;
}

template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_taper_all(T *__restrict__ P0, T *__restrict__ P1, const float *__restrict__ t, Grid g,
                int b, int nA, int nB, TaperDivs dv)
{
    T *__restrict__ P = blockIdx.y ? P1 : P0;
    const int i0 = (int)blockIdx.x * (kThreads * kTaperPer) + (int)threadIdx.x;
    TaperCell c[kTaperPer];
    float v[kTaperPer];
#pragma unroll
    for (int k = 0; k < kTaperPer; ++k) {
        c[k] = taper_cell(t, g, b, nA, nB, dv, i0 + k * kThreads);
        v[k] = c[k].ok ? wld(P[c[k].off]) : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < kTaperPer; ++k) {
        if (!c[k].ok) continue;
        float w = v[k] * c[k].f1;
        w *= c[k].f2;
        w *= c[k].f3;
        P[c[k].off] = wst<T>(w);
    }
}

// Snapshot of a wavefield over the imaging range (all x, y; z in [b, nz-b)):
// snap is [nx][ny][nz-2b] of element type S (float, or _Float16 rounded to
// nearest-even).  One wave per (x, y) row, z on the lanes, V consecutive
// cells per lane (V = 4 when every row start and both arrays allow the
// vector access: one 16- or 8-byte load and one 8- or 16-byte store per
// lane instead of 2-byte accesses; the values are the same).
template <typename E, int N>
struct alignas(sizeof(E) * N) SnapVec {
    E v[N];
};

template <typename T, typename S, int V>
__global__ void __launch_bounds__(kThreads)
    k_snap_store(const T *__restrict__ P, S *__restrict__ snap, Grid g, int b)
{
    const int lane = threadIdx.x & 63;
    const int row = __builtin_amdgcn_readfirstlane((int)blockIdx.x * (kThreads / 64) +
                                                   (int)(threadIdx.x >> 6));
    if (row >= g.nx * g.ny) return;
    const int x = row / g.ny;
    const int y = row - x * g.ny;
    const int lz = g.nz - 2 * b;
    const T *__restrict__ r = P + ((long long)x * g.sx + (long long)y * g.sy + b);
    S *__restrict__ o = snap + (long long)row * lz;
    if constexpr (V == 1) {
        for (int z = lane; z < lz; z += 64) o[z] = wst<S>(wld(r[z]));
    } else {
        const SnapVec<T, V> *__restrict__ rv = reinterpret_cast<const SnapVec<T, V> *>(r);
        SnapVec<S, V> *__restrict__ ov = reinterpret_cast<SnapVec<S, V> *>(o);
        for (int k = lane; k < lz / V; k += 64) {
            const SnapVec<T, V> a = rv[k];
            SnapVec<S, V> c;
#pragma unroll
            for (int e = 0; e < V; ++e) c.v[e] = wst<S>(wld(a.v[e]));
            ov[k] = c;
        }
    }
}

// Imaging condition from a snapshot: img += w * (snap * PR) on the imaging
// range (the reference's img += PS * PR, RTMCPUPlatform.cpp:513-540, with PS
// read from the snapshot and weighted by the snapshot interval w; w = 1
// leaves the sum bit-identical to img + PS * PR).
template <typename T, typename S>
__global__ void __launch_bounds__(kThreads)
    k_snap_image(float *__restrict__ img, const S *__restrict__ snap, const T *__restrict__ PR,
                 Grid g, int b, float w)
{
    const int lane = threadIdx.x & 63;
    const int row = __builtin_amdgcn_readfirstlane((int)blockIdx.x * (kThreads / 64) +
                                                   (int)(threadIdx.x >> 6));
    if (row >= g.nx * g.ny) return;
    const int x = row / g.ny;
    const int y = row - x * g.ny;
    const int lz = g.nz - 2 * b;
    const long long o = (long long)x * g.sx + (long long)y * g.sy + b;
    const S *__restrict__ sr = snap + (long long)row * lz;
    for (int z = lane; z < lz; z += 64) {
        const float v = wld(sr[z]) * wld(PR[o + z]);
        img[o + z] = img[o + z] + w * v;
    }
}

// rtmApplySource (RTMCPUPlatform.cpp:103-113): PP[s] += value (or table[it]).
template <typename T>
__global__ void k_source(T *__restrict__ PP, long long o, const float *__restrict__ table, int it,
                         float value)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const float v = table ? table[it] : value;
        PP[o] = wst<T>(wld(PP[o]) + v);
    }
}

// rtmSaveReceiverData / rtmRestoreReceiverData (RTMCPUPlatform.cpp:114-151).
template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_receivers(T *__restrict__ PP, float *__restrict__ rcv, Grid g, int rx0, int ry0, int rz,
                int rdx, int rdy, int rnx, int rny, int ntstep, int lt, int restore)
{
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= rnx * rny) return;
    const int ix = i / rny, iy = i % rny;
    const int rx = rx0 + ix * rdx, ry = ry0 + iy * rdy;
    if (rx < 0 || rx >= g.nx || ry < 0 || ry >= g.ny) return;
    const long long o = goff(g, rx, ry, rz);
    const long long ro = (long long)i * ntstep + lt;
    if (restore) PP[o] = wst<T>(rcv[ro]);
    else rcv[ro] = wld(PP[o]);
}

// Epilogue of the fused backward receiver step.  Threads [0, rnx*rny):
// receiver injection (SET) followed by the imaging condition at the receiver
// cells.  Remaining threads: the imaging condition on the x/y frame of width
// R that the stencil kernel does not visit (all x, y are imaged by the
// reference, RTMCPUPlatform.cpp:513-540), skipping receiver cells.
__device__ __forceinline__ bool is_receiver(int x, int y, int z, int rx0, int ry0, int rz,
                                            int rdx, int rdy, int rnx, int rny)
{
    const int dx = x - rx0, dy = y - ry0;
    return z == rz && dx >= 0 && dy >= 0 && (dx % rdx) == 0 && (dy % rdy) == 0 &&
           (dx / rdx) < rnx && (dy / rdy) < rny;
}

// The x / y frame of width R on z in [z0, z1): the cells of the imaging
// range that the stencil kernel does not visit.
__device__ __forceinline__ long long image_frame_count(const Grid &g, int R, int z0, int z1)
{
    const long long lz = z1 > z0 ? z1 - z0 : 0;
    return (long long)2 * R * g.ny * lz + (long long)(g.nx - 2 * R) * 2 * R * lz;
}
template <typename I>
__device__ __forceinline__ bool image_frame_cell(const Grid &g, int R, int z0, int z1, I i,
                                                 int &x, int &y, int &z)
{
    const int lz = z1 - z0;
    if (lz <= 0) return false;
    const I nA = (I)2 * R * g.ny * lz;
    const I nB = (I)(g.nx - 2 * R) * 2 * R * lz;
    if (i < nA) {
        z = z0 + (int)(i % lz);
        const I r = i / lz;
        y = (int)(r % g.ny);
        const int xx = (int)(r / g.ny);
        x = xx < R ? xx : g.nx - 2 * R + xx;
    } else if ((i -= nA) < nB) {
        z = z0 + (int)(i % lz);
        const I r = i / lz;
        const int yy = (int)(r % (2 * R));
        x = R + (int)(r / (2 * R));
        y = yy < R ? yy : g.ny - 2 * R + yy;
    } else {
        return false;
    }
    return true;
}
// Receiver i of the grid: injection (SET) of v into PP, then the imaging
// condition at that cell with the rounded value.
template <typename T>
__device__ __forceinline__ void inject_image(T *__restrict__ PP, float *__restrict__ img,
                                             const T *__restrict__ ps, const Grid &g, int rx,
                                             int ry, int rz, float v, int z0, int z1)
{
    if (rx < 0 || rx >= g.nx || ry < 0 || ry >= g.ny) return;
    const long long o = goff(g, rx, ry, rz);
    PP[o] = wst<T>(v);
    if (rz >= z0 && rz < z1) img[o] = img[o] + wld(ps[o]) * wrnd<T>(v);
}

template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_receivers_image(T *__restrict__ PP, const float *__restrict__ rcv, Grid g, int rx0, int ry0,
                      int rz, int rdx, int rdy, int rnx, int rny, int ntstep, int lt,
                      float *__restrict__ img, const T *__restrict__ ps, int z0, int z1, int R,
                      FrameDivs dv, FastDiv drny)
{
    const int i = (int)(blockIdx.x * kThreads + threadIdx.x);  // < 2^31 (launcher)
    const int nrcv = rnx * rny;
    if (i < nrcv) {
        const int ix = fdiv(i, drny), iy = i - ix * rny;
        inject_image(PP, img, ps, g, rx0 + ix * rdx, ry0 + iy * rdy, rz,
                     rcv[(long long)i * ntstep + lt], z0, z1);
        return;
    }
    int x, y, z;
    if (!image_frame_cell_fd(g, R, z0, z1, i - nrcv, dv, x, y, z)) return;
    if (z == rz && is_receiver(x, y, z, rx0, ry0, rz, rdx, rdy, rnx, rny)) return;
    const long long o = goff(g, x, y, z);
    img[o] = img[o] + wld(ps[o]) * wld(PP[o]);
}

// rtmSaveUpperBorder / rtmRestoreUpperBorder (RTMCPUPlatform.cpp:153-189):
// upb[lt][x][y][k] <-> PP[x][y][blen-hf+k], 64-bit offsets.
// mode 0 save, 1 restore, 2 save and then apply the upper taper (table t)
// to the saved cell in place.
// 32-bit cell index (n = nx * ny * hf < 2^31, launcher) with divisions by
// launch constants through a multiply-high; the history offset is 64-bit.
template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_upb(T *__restrict__ PP, T *__restrict__ upb, Grid g, int b, int hf, long long lt, int mode,
          const float *__restrict__ t, int n, FastDiv dhf, FastDiv dny)
{
    const int i = (int)(blockIdx.x * kThreads + threadIdx.x);
    if (i >= n) return;
    const int r = fdiv(i, dhf);
    const int k = i - r * hf;
    const int x = fdiv(r, dny);
    const int y = r - x * g.ny;
    const long long u = lt * (long long)n + i;
    const int z = b - hf + k;
    const long long o = goff(g, x, y, z);
    if (mode == 1) {
        PP[o] = upb[u];
    } else {
        const T v = PP[o];
        upb[u] = v;
        if (mode == 2) PP[o] = wst<T>(frame_taper<T>(wld(v), t, g, b, x, y, z, 1));
    }
}

// rtmImageCondition (RTMCPUPlatform.cpp:513-540): z in [z0, z1), all x, y.
template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_image(float *__restrict__ img, const T *__restrict__ PS, const T *__restrict__ PR, Grid g,
            int z0, int z1)
{
    const int lz = z1 - z0;
    const long long n = (long long)g.nx * g.ny * lz;
    const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const int z = z0 + (int)(i % lz);
    const long long r = i / lz;
    const int y = (int)(r % g.ny);
    const int x = (int)(r / g.ny);
    const long long o = goff(g, x, y, z);
    img[o] = img[o] + wld(PS[o]) * wld(PR[o]);
}

// Order-2 Laplacian filter (RTMGrid.cpp:226-300) evaluated only where
// removeBorders (RTMGrid.cpp:129-166) keeps it; out is dense [nx][ny][nz].
__global__ void __launch_bounds__(kThreads)
    k_image_finish(const float *__restrict__ img, float *__restrict__ out, Grid g, int b,
                   Coefs fc)
{
    const int nx = g.nx - 2 * b, ny = g.ny - 2 * b, nz = g.nz - 2 * b;
    const long long n = (long long)nx * ny * nz;
    const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const int z = (int)(i % nz);
    const long long r = i / nz;
    const int y = (int)(r % ny);
    const int x = (int)(r / ny);
    const long long o = goff(g, x + b, y + b, z + b);
    float ax = 0.f, ay = 0.f, az = 0.f;
#pragma unroll
    for (int io = 0; io < 3; ++io) {
        ax += img[o + (io - 1) * g.sx] * fc.c[0][io];
        ay += img[o + (long long)(io - 1) * g.sy] * fc.c[1][io];
        az += img[o + (io - 1)] * fc.c[2][io];
    }
    out[i] = ax + ay + az;
}

// rtmUpdateFreqContributions then rtmFreqDomainImageCondition
// (RTMCPUPlatform.cpp:20-96) for the frequencies iw = iw0 + lw (lw < nwstep,
// iw < nw) at time index st, one thread per interior cell (extended offset
// +b), frequencies in order: sRe += kre*ps, sIm += kim*ps, rRe += kre*pr,
// rIm += kim*pr, img += w2*(sRe*rRe - sIm*rIm).  Accumulators are
// [nwstep][nx][ny][nz]; the image stays in a register across the group.
__global__ void __launch_bounds__(kThreads)
    k_freq_image(float *__restrict__ img, const float *__restrict__ PS,
                 const float *__restrict__ PR, float *__restrict__ sre, float *__restrict__ sim,
                 float *__restrict__ rre, float *__restrict__ rim, const float *__restrict__ kre,
                 const float *__restrict__ kim, const float *__restrict__ w2, Grid g, int b,
                 int nx, int ny, int nz, int nwstep, int iw0, int nw, int nt, int st)
{
    const long long vol = (long long)nx * ny * nz;
    const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (i >= vol) return;
    const int iz = (int)(i % nz);
    const long long r = i / nz;
    const int iy = (int)(r % ny), ix = (int)(r / ny);
    const long long o = goff(g, ix + b, iy + b, iz + b);
    const float ps = PS[o], pr = PR[o];
    float im = img[o];
    for (int lw = 0, iw = iw0; lw < nwstep && iw < nw; ++lw, ++iw) {
        const float kr = kre[(long long)iw * nt + st], ki = kim[(long long)iw * nt + st];
        const long long f = lw * vol + i;
        const float a = sre[f] + kr * ps;
        const float c2 = sim[f] + ki * ps;
        const float c = rre[f] + kr * pr;
        const float d = rim[f] + ki * pr;
        sre[f] = a;
        sim[f] = c2;
        rre[f] = c;
        rim[f] = d;
        im = im + w2[iw] * (a * c - c2 * d);
    }
    img[o] = im;
}

__global__ void __launch_bounds__(kThreads) k_fill(float *__restrict__ p, size_t n, float v)
{
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n;
         i += (size_t)gridDim.x * kThreads)
        p[i] = v;
}

__global__ void __launch_bounds__(kThreads)
    k_fill_u16(uint16_t *__restrict__ p, size_t n, uint16_t v)
{
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n;
         i += (size_t)gridDim.x * kThreads)
        p[i] = v;
}

__global__ void __launch_bounds__(kThreads)
    k_stack(float *__restrict__ d, const float *__restrict__ s, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n;
         i += (size_t)gridDim.x * kThreads)
        d[i] = d[i] + s[i];
}

// Batched tails: one launch for every shot of a batch.  Each shot's work
// items fill whole blocks, so a block finds its shot (a binary search over
// the shots' first blocks) with block-uniform scalar loads; per-shot item
// counts stay below 2^31 (checked when the batch is made).
template <int BatchShot::*first>
__device__ __forceinline__ int batch_shot(const BatchShot *__restrict__ shots, int n, int blk)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (shots[mid].*first <= blk) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Forward tail of an HBC step on PP = fs[par ^ 1]: the frame taper, the
// source injection and the UPB save of step `it`, in that order per cell,
// as launch_taper_frame, launch_source and launch_upb would run them one
// after the other.  Work items per shot: the UPB layer (z in [b-hf, b),
// frame cells among them tapered first), the frame cells outside the layer,
// and the source cell when it lies in neither.
template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_tail_fwd(const BatchShot *__restrict__ shots, int n, int par, const float *__restrict__ t,
               int b, int R, int hf, int taps, const int *__restrict__ itp, int itoff, int out)
{
    const BatchShot &sh = shots[batch_shot<&BatchShot::fwd0>(shots, n, (int)blockIdx.x)];
    int i = ((int)blockIdx.x - sh.fwd0) * kThreads + (int)threadIdx.x;
    if (i >= sh.nfwd) return;
    const Grid g = sh.g;
    const int it = (itp ? *itp : 0) + itoff;
    T *PP = static_cast<T *>(sh.fs[par ^ 1]);
    const int nl = g.nx * g.ny * hf;
    int x, y, z;
    bool save = false;
    if (i < nl) {
        const int r = i / hf;
        z = b - hf + (int)(i % hf);
        y = (int)(r % g.ny);
        x = (int)(r / g.ny);
        save = true;
    } else {
        i -= nl;
        const int nf = (int)frame_count(g, b, R);
        if (i < nf) {
            frame_cell(g, b, R, i, x, y, z);
            if (z >= b - hf) return;  // a layer cell: its layer item handles it
        } else {
            const long long so = sh.src_off;
            x = (int)(so / g.sx);
            y = (int)((so % g.sx) / g.sy);
            z = (int)(so % g.sy);
            if ((z >= b - hf && z < b) || in_frame(g, b, R, x, y, z)) return;
        }
    }
    const long long o = goff(g, x, y, z);
    const bool fr = in_frame(g, b, R, x, y, z), src = o == sh.src_off;
    if (!fr && !src && !save) return;
    T tv = PP[o];
    // out: the step stored the upper taper of its output except in the
    // layer, which is saved first and then tapered once (its frame cells
    // carry one taper fewer until then)
    const int ftaps = (out && save) ? taps - 1 : taps;
    if (fr || src) {
        float v = wld(tv);
        if (fr) v = wrnd<T>(frame_taper<T>(v, t, g, b, x, y, z, ftaps));
        if (src) v = v + sh.ricker[it];
        tv = wst<T>(v);
        PP[o] = tv;
    }
    if (save) {
        static_cast<T *>(sh.upb)[(long long)it * nl + i] = tv;
        if (out) PP[o] = wst<T>(frame_taper<T>(wld(tv), t, g, b, x, y, z, 1));
    }
}

// Backward tail of an HBC step: UPB restore of step `it` into the new source
// field PPs = fs[spar ^ 1]; frame taper, receiver injection (SET) and the
// imaging condition off the stencil's cells on PPr = fr[rpar ^ 1] with the
// source value ps = PPs.  The four cell sets are disjoint (the layer and the
// frame lie above z = b, receivers and the imaging range at or below it), so
// one launch reproduces launch_upb, launch_taper_frame and
// launch_receivers_image run in sequence.
template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_tail_bwd(const BatchShot *__restrict__ shots, int n, int spar, int rpar,
               const float *__restrict__ t, int b, int R, int hf, int taps, int nt,
               const int *__restrict__ itp, int itoff)
{
    const BatchShot &sh = shots[batch_shot<&BatchShot::bwd0>(shots, n, (int)blockIdx.x)];
    int i = ((int)blockIdx.x - sh.bwd0) * kThreads + (int)threadIdx.x;
    if (i >= sh.nbwd) return;
    const Grid g = sh.g;
    const int it = (itp ? *itp : 0) + itoff;
    T *PPs = static_cast<T *>(sh.fs[spar ^ 1]);
    T *PPr = static_cast<T *>(sh.fr[rpar ^ 1]);
    const int nl = g.nx * g.ny * hf;
    if (i < nl) {
        const int r = i / hf;
        const int k = (int)(i % hf);
        PPs[goff(g, (int)(r / g.ny), (int)(r % g.ny), b - hf + k)] =
            static_cast<const T *>(sh.upb)[(long long)it * nl + i];
        return;
    }
    i -= nl;
    int x, y, z;
    const int nf = (int)frame_count(g, b, R);
    if (i < nf) {
        frame_cell(g, b, R, i, x, y, z);
        const long long o = goff(g, x, y, z);
        PPr[o] = wst<T>(frame_taper<T>(wld(PPr[o]), t, g, b, x, y, z, taps));
        return;
    }
    i -= nf;
    const int z0 = b, z1 = g.nz - b;
    const int nrcv = sh.rnx * sh.rny;
    if (i < nrcv) {
        const int ix = (int)(i / sh.rny), iy = (int)(i % sh.rny);
        inject_image(PPr, sh.img, (const T *)PPs, g, sh.rx0 + ix * sh.rdx, sh.ry0 + iy * sh.rdy,
                     sh.rz, sh.rcv[(long long)i * nt + it], z0, z1);
        return;
    }
    if (!image_frame_cell(g, R, z0, z1, i - nrcv, x, y, z)) return;
    if (is_receiver(x, y, z, sh.rx0, sh.ry0, sh.rz, sh.rdx, sh.rdy, sh.rnx, sh.rny)) return;
    const long long o = goff(g, x, y, z);
    sh.img[o] = sh.img[o] + wld(PPs[o]) * wld(PPr[o]);
}

__global__ void k_tick(int *p, int d, int set)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = set ? d : *p + d;
}

// The 16-byte group kernels apply: G-cell groups never straddle a border
// edge or a row, and every group is 16-byte aligned.
template <typename T>
bool vec16_ok(const T *P, const T *Q, const Grid &g, int b, int G)
{
    const auto al = [](const void *p) { return ((uintptr_t)p & 15u) == 0; };
    return b % G == 0 && g.nz % G == 0 && g.sy % G == 0 && g.sx % G == 0 && al(P) &&
           (!Q || al(Q));
}

inline unsigned nblocks(long long n)
{
    return (unsigned)((n + kThreads - 1) / kThreads);
}

inline unsigned stride_blocks(size_t n)
{
    size_t b = (n + kThreads - 1) / kThreads;
    return (unsigned)(b > 8192 ? 8192 : (b ? b : 1));
}

}  // namespace

template <typename T>
hipError_t launch_taper(T *P, const float *taper, const Grid &g, int b, bool upper_only,
                        hipStream_t s)
{
    if (b <= 0) return hipSuccess;
    if (upper_only) return launch_taper_upper2(P, (T *)nullptr, taper, g, b, s);
    return launch_taper_all2(P, (T *)nullptr, taper, g, b, s);
}

template <typename T>
hipError_t launch_taper_upper2(T *P, T *Q, const float *taper, const Grid &g, int b,
                               hipStream_t s)
{
    if (b <= 0) return hipSuccess;
    const long long n = (long long)g.nx * g.ny * b;
    const long long per = (long long)kThreads * kUpperPer;
    if (n + per >= (1LL << 31)) return hipErrorInvalidValue;
    constexpr int G = Group16<T>::G;
    if (vec16_ok(P, Q, g, b, G)) {
        const long long ng = n / G;
        hipLaunchKernelGGL(k_taper_upper_v<T>, dim3(nblocks(ng), Q ? 2 : 1), dim3(kThreads), 0, s,
                           P, Q, taper, g, b, (int)ng, fast_div(b / G), fast_div(g.ny));
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_taper_upper<T>, dim3((unsigned)((n + per - 1) / per), Q ? 2 : 1),
                       dim3(kThreads), 0, s, P, Q, taper, g, b, (int)n, fast_div(b),
                       fast_div(g.ny));
    return hipGetLastError();
}

template <typename T>
hipError_t launch_taper_all2(T *P, T *Q, const float *taper, const Grid &g, int b,
                             hipStream_t s)
{
    if (b <= 0) return hipSuccess;
    if ((long long)g.nx * g.ny >= (1LL << 31)) return hipErrorInvalidValue;
    const unsigned ng = Q ? 2 : 1;
    // the front-right edge first: it reads (blen, y, z) before the face pass
    // rescales that cell
    const long long lz = g.nz - 2 * b;
    const long long nfr = (long long)b * b * lz;
    if (nfr > 0 && g.nx > 2 * b) {
        hipLaunchKernelGGL(k_taper_front_right<T>, dim3(nblocks(nfr), ng), dim3(kThreads), 0, s,
                           P, Q, taper, g, b);
    }
    const long long nA = (long long)(g.nx - 2 * b) * (g.ny - 2 * b) * 2 * b;
    const long long nB = ((long long)2 * b * g.ny + (long long)(g.nx - 2 * b) * 2 * b) * g.nz;
    const long long per = (long long)kThreads * kTaperPer;
    if (nA + nB + per >= (1LL << 31)) return hipErrorInvalidValue;
    constexpr int G = Group16<T>::G;
    if (vec16_ok(P, Q, g, b, G)) {
        const long long ngr = (nA + nB) / G;
        hipLaunchKernelGGL(k_taper_all_v<T>, dim3(nblocks(ngr), ng), dim3(kThreads), 0, s, P, Q,
                           taper, g, b, (int)(nA / G), (int)(nB / G), fast_div(2 * b / G),
                           fast_div(g.ny - 2 * b), fast_div(g.nz / G), fast_div(g.ny),
                           fast_div(2 * b));
        return hipGetLastError();
    }
    const TaperDivs dv{fast_div(2 * b), fast_div(g.ny - 2 * b), fast_div(g.nz), fast_div(g.ny)};
    hipLaunchKernelGGL(k_taper_all<T>, dim3((unsigned)((nA + nB + per - 1) / per), ng),
                       dim3(kThreads), 0, s, P, Q, taper, g, b, (int)nA, (int)nB, dv);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_taper_frame(T *P, const float *taper, const Grid &g, int b, int R, int taps,
                              hipStream_t s, int hold)
{
    if (b <= 0 || taps <= 0) return hipSuccess;
    if (hold < 0 || hold > b) return hipErrorInvalidValue;
    const long long lz = b - R;
    const long long n = (long long)g.nx * g.ny * R + (long long)2 * R * g.ny * lz +
                        (long long)(g.nx - 2 * R) * 2 * R * lz;
    if (n + kThreads >= (1LL << 31)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_taper_frame<T>, dim3(nblocks(n)), dim3(kThreads), 0, s, P, taper, g, b, R,
                       taps, frame_divs(R, b - R, g.ny), b - hold);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_taper_frame_all(T *P, const float *taper, const Grid &g, int b, int R, int taps,
                                  hipStream_t s)
{
    if (b <= 0 || taps <= 0) return hipSuccess;
    if (b < R || 2 * b >= g.nx || 2 * b >= g.ny || 2 * b >= g.nz) return hipErrorInvalidValue;
    const long long nX = (long long)2 * R * g.ny * g.nz;
    const long long n = nX + (long long)(g.nx - 2 * R) * 2 * R * g.nz;
    const long long lz = g.nz - 2 * b;
    const long long n1 = (long long)R * b * lz;
    const long long nq = n1 + (long long)(b - R) * R * lz;
    if (n + kThreads >= (1LL << 31)) return hipErrorInvalidValue;
    if (nq > 0) {
        hipLaunchKernelGGL(k_taper_frame_quirk<T>, dim3(nblocks(nq)), dim3(kThreads), 0, s, P,
                           taper, g, b, R, taps, (int)n1, (int)nq);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_taper_frame_all<T>, dim3(nblocks(n)), dim3(kThreads), 0, s, P, taper, g,
                       b, R, taps, (int)nX, (int)n);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_taper_frame_xy(T *P, const float *taper, const Grid &g, int b, int R, int taps,
                                 hipStream_t s)
{
    if (b <= 0 || taps <= 0) return hipSuccess;
    if (b < R || 2 * b >= g.nx || 2 * b >= g.ny || 2 * b >= g.nz) return hipErrorInvalidValue;
    const long long nX = (long long)2 * R * g.ny * g.nz;
    const long long n = nX + (long long)(g.nx - 2 * R) * 2 * R * g.nz;
    if (n + kThreads >= (1LL << 31)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_taper_frame_all<T>, dim3(nblocks(n)), dim3(kThreads), 0, s, P, taper, g,
                       b, R, taps, (int)nX, (int)n);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_taper_quirk_out(T *P, const float *taper, const Grid &g, int b, hipStream_t s)
{
    if (b <= 0) return hipSuccess;
    if (2 * b >= g.nx || 2 * b >= g.ny || 2 * b >= g.nz) return hipErrorInvalidValue;
    const long long lz = g.nz - 2 * b;
    const long long n = (long long)b * b * lz;
    if (n + kThreads >= (1LL << 31)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_taper_quirk_out<T>, dim3(nblocks(n)), dim3(kThreads), 0, s, P, taper, g,
                       b, fast_div((uint32_t)lz), fast_div((uint32_t)b), (int)n);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_source(T *PP, long long offset, const float *table, int it, float value,
                         hipStream_t s)
{
    hipLaunchKernelGGL(k_source<T>, dim3(1), dim3(64), 0, s, PP, offset, table, it, value);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_receivers(T *PP, float *rcv, const Grid &g, int rx0, int ry0, int rz,
                            int rdx, int rdy, int rnx, int rny, int ntstep, int lt, bool restore,
                            hipStream_t s)
{
    const long long n = (long long)rnx * rny;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_receivers<T>, dim3(nblocks(n)), dim3(kThreads), 0, s, PP, rcv, g, rx0, ry0,
                       rz, rdx, rdy, rnx, rny, ntstep, lt, restore ? 1 : 0);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_receivers_image(T *PP, const float *rcv, const Grid &g, int rx0, int ry0,
                                  int rz, int rdx, int rdy, int rnx, int rny, int ntstep, int lt,
                                  float *img, const T *ps, int z0, int z1, int R,
                                  hipStream_t s)
{
    const long long lz = z1 > z0 ? z1 - z0 : 0;
    const long long n = (long long)rnx * rny + (long long)2 * R * g.ny * lz +
                        (long long)(g.nx - 2 * R) * 2 * R * lz;
    if (n <= 0) return hipSuccess;
    if (n + kThreads >= (1LL << 31)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_receivers_image<T>, dim3(nblocks(n)), dim3(kThreads), 0, s, PP, rcv, g, rx0,
                       ry0, rz, rdx < 1 ? 1 : rdx, rdy < 1 ? 1 : rdy, rnx, rny, ntstep, lt, img,
                       ps, z0, z1, R, frame_divs(R, (int)lz, g.ny),
                       fast_div((uint32_t)std::max(rny, 1)));
    return hipGetLastError();
}

template <typename T>
hipError_t launch_upb(T *PP, T *upb, const Grid &g, int b, int hf, long long lt,
                      bool restore, hipStream_t s)
{
    const long long n = (long long)g.nx * g.ny * hf;
    if (hf <= 0 || n + kThreads >= (1LL << 31)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_upb<T>, dim3(nblocks(n)), dim3(kThreads), 0, s, PP, upb, g, b, hf, lt,
                       restore ? 1 : 0, (const float *)nullptr, (int)n, fast_div((uint32_t)hf),
                       fast_div((uint32_t)g.ny));
    return hipGetLastError();
}

template <typename T>
hipError_t launch_upb_save_taper(T *PP, T *upb, const float *taper, const Grid &g, int b, int hf,
                                 long long lt, hipStream_t s)
{
    if (!taper || hf <= 0 || hf > b) return hipErrorInvalidValue;
    const long long n = (long long)g.nx * g.ny * hf;
    if (n + kThreads >= (1LL << 31)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_upb<T>, dim3(nblocks(n)), dim3(kThreads), 0, s, PP, upb, g, b, hf, lt, 2,
                       taper, (int)n, fast_div((uint32_t)hf), fast_div((uint32_t)g.ny));
    return hipGetLastError();
}

template <typename T>
hipError_t launch_image(float *img, const T *PS, const T *PR, const Grid &g, int z0,
                        int z1, hipStream_t s)
{
    if (z1 <= z0) return hipSuccess;
    const long long n = (long long)g.nx * g.ny * (z1 - z0);
    hipLaunchKernelGGL(k_image<T>, dim3(nblocks(n)), dim3(kThreads), 0, s, img, PS, PR, g, z0, z1);
    return hipGetLastError();
}

template <typename T, typename S>
hipError_t launch_snap_store(const T *P, S *snap, const Grid &g, int b, hipStream_t s)
{
    if (g.nz <= 2 * b) return hipSuccess;
    if ((long long)g.nx * g.ny >= (1LL << 31)) return hipErrorInvalidValue;
    const long long rows = (long long)g.nx * g.ny;
    const dim3 grid((unsigned)((rows + 3) / 4));
    // vector form: every row start (x*sx + y*sy + b in P, row*(nz-2b) in
    // snap) a multiple of 4 cells and both arrays aligned to 4 cells
    const bool vec = (b % 4) == 0 && (g.sy % 4) == 0 && (g.sx % 4) == 0 &&
                     ((g.nz - 2 * b) % 4) == 0 && ((uintptr_t)P % (4 * sizeof(T))) == 0 &&
                     ((uintptr_t)snap % (4 * sizeof(S))) == 0;
    if (vec)
        hipLaunchKernelGGL((k_snap_store<T, S, 4>), grid, dim3(kThreads), 0, s, P, snap, g, b);
    else
        hipLaunchKernelGGL((k_snap_store<T, S, 1>), grid, dim3(kThreads), 0, s, P, snap, g, b);
    return hipGetLastError();
}

template <typename T, typename S>
hipError_t launch_snap_image(float *img, const S *snap, const T *PR, const Grid &g, int b,
                             float w, hipStream_t s)
{
    if (g.nz <= 2 * b) return hipSuccess;
    if ((long long)g.nx * g.ny >= (1LL << 31)) return hipErrorInvalidValue;
    const long long rows = (long long)g.nx * g.ny;
    hipLaunchKernelGGL((k_snap_image<T, S>), dim3((unsigned)((rows + 3) / 4)), dim3(kThreads), 0,
                       s, img, snap, PR, g, b, w);
    return hipGetLastError();
}

hipError_t launch_image_finish(const float *img, float *out, const Grid &g, int b,
                               const Coefs &fc, hipStream_t s)
{
    const long long n = (long long)(g.nx - 2 * b) * (g.ny - 2 * b) * (g.nz - 2 * b);
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_image_finish, dim3(nblocks(n)), dim3(kThreads), 0, s, img, out, g, b, fc);
    return hipGetLastError();
}

hipError_t launch_freq_image(float *img, const float *PS, const float *PR, float *sre,
                             float *sim, float *rre, float *rim, const float *kre,
                             const float *kim, const float *w2, const Grid &g, int b, int nwstep,
                             int iw0, int nw, int nt, int st, hipStream_t s)
{
    const int nx = g.nx - 2 * b, ny = g.ny - 2 * b, nz = g.nz - 2 * b;
    const long long vol = (long long)nx * ny * nz;
    if (vol <= 0 || nwstep <= 0 || iw0 >= nw) return hipSuccess;
    hipLaunchKernelGGL(k_freq_image, dim3(nblocks(vol)), dim3(kThreads), 0, s, img, PS, PR, sre,
                       sim, rre, rim, kre, kim, w2, g, b, nx, ny, nz, nwstep, iw0, nw, nt, st);
    return hipGetLastError();
}

hipError_t launch_fill(float *p, size_t n, float v, hipStream_t s)
{
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_fill, dim3(stride_blocks(n)), dim3(kThreads), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launch_fill_u16(uint16_t *p, size_t n, uint16_t bits, hipStream_t s)
{
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u16, dim3(stride_blocks(n)), dim3(kThreads), 0, s, p, n, bits);
    return hipGetLastError();
}

hipError_t launch_stack(float *dst, const float *src, size_t n, hipStream_t s)
{
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_stack, dim3(stride_blocks(n)), dim3(kThreads), 0, s, dst, src, n);
    return hipGetLastError();
}

long long tail_fwd_items(const Grid &g, int b, int R, int hf)
{
    const long long lz = b - R;
    return (long long)g.nx * g.ny * hf + (long long)g.nx * g.ny * R +
           (long long)2 * R * g.ny * lz + (long long)(g.nx - 2 * R) * 2 * R * lz + 1;
}

long long tail_bwd_items(const Grid &g, int b, int R, int hf, int rnx, int rny)
{
    const long long lz = b - R, li = g.nz - 2 * b > 0 ? g.nz - 2 * b : 0;
    return (long long)g.nx * g.ny * hf + (long long)g.nx * g.ny * R +
           (long long)2 * R * g.ny * lz + (long long)(g.nx - 2 * R) * 2 * R * lz +
           (long long)rnx * rny + (long long)2 * R * g.ny * li +
           (long long)(g.nx - 2 * R) * 2 * R * li;
}

int tail_blocks(long long items) { return (int)((items + kThreads - 1) / kThreads); }

template <typename T>
hipError_t launch_tail_fwd(const BatchShot *shots, int n, int blocks, int par,
                           const float *taper, int b, int R, int hf, int taps, const int *itp,
                           int itoff, hipStream_t s, bool out)
{
    if (blocks <= 0 || n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_tail_fwd<T>, dim3(blocks), dim3(kThreads), 0, s, shots, n, par, taper, b,
                       R, hf, taps, itp, itoff, out ? 1 : 0);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_tail_bwd(const BatchShot *shots, int n, int blocks, int spar, int rpar,
                           const float *taper, int b, int R, int hf, int taps, int nt,
                           const int *itp, int itoff, hipStream_t s)
{
    if (blocks <= 0 || n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_tail_bwd<T>, dim3(blocks), dim3(kThreads), 0, s, shots, n, spar, rpar,
                       taper, b, R, hf, taps, nt, itp, itoff);
    return hipGetLastError();
}

hipError_t launch_tick(int *p, int d, bool set, hipStream_t s)
{
    hipLaunchKernelGGL(k_tick, dim3(1), dim3(64), 0, s, p, d, set ? 1 : 0);
    return hipGetLastError();
}

#define RTM_INSTANTIATE_OPS(T)                                                                \
    template hipError_t launch_taper<T>(T *, const float *, const Grid &, int, bool,            \
                                        hipStream_t);                                           \
    template hipError_t launch_taper_all2<T>(T *, T *, const float *, const Grid &, int,        \
                                             hipStream_t);                                      \
    template hipError_t launch_taper_upper2<T>(T *, T *, const float *, const Grid &, int,      \
                                               hipStream_t);                                    \
    template hipError_t launch_taper_frame<T>(T *, const float *, const Grid &, int, int, int,  \
                                              hipStream_t, int);                                \
    template hipError_t launch_taper_frame_all<T>(T *, const float *, const Grid &, int, int,   \
                                                  int, hipStream_t);                            \
    template hipError_t launch_taper_frame_xy<T>(T *, const float *, const Grid &, int, int,    \
                                                 int, hipStream_t);                             \
    template hipError_t launch_taper_quirk_out<T>(T *, const float *, const Grid &, int,        \
                                                  hipStream_t);                                 \
    template hipError_t launch_source<T>(T *, long long, const float *, int, float, hipStream_t); \
    template hipError_t launch_receivers<T>(T *, float *, const Grid &, int, int, int, int, int, \
                                            int, int, int, int, bool, hipStream_t);             \
    template hipError_t launch_receivers_image<T>(T *, const float *, const Grid &, int, int,   \
                                                  int, int, int, int, int, int, int, float *,   \
                                                  const T *, int, int, int, hipStream_t);       \
    template hipError_t launch_upb<T>(T *, T *, const Grid &, int, int, long long, bool,        \
                                      hipStream_t);                                             \
    template hipError_t launch_upb_save_taper<T>(T *, T *, const float *, const Grid &, int,    \
                                                 int, long long, hipStream_t);                  \
    template hipError_t launch_image<T>(float *, const T *, const T *, const Grid &, int, int,  \
                                        hipStream_t);                                           \
    template hipError_t launch_tail_fwd<T>(const BatchShot *, int, int, int, const float *,    \
                                           int, int, int, int, const int *, int, hipStream_t,   \
                                           bool);                                               \
    template hipError_t launch_tail_bwd<T>(const BatchShot *, int, int, int, int,               \
                                           const float *, int, int, int, int, int, const int *, \
                                           int, hipStream_t);
RTM_INSTANTIATE_OPS(float)
RTM_INSTANTIATE_OPS(_Float16)
RTM_INSTANTIATE_OPS(__bf16)

#define RTM_INSTANTIATE_SNAP(T, S)                                                            \
    template hipError_t launch_snap_store<T, S>(const T *, S *, const Grid &, int, hipStream_t); \
    template hipError_t launch_snap_image<T, S>(float *, const S *, const T *, const Grid &, int, \
                                                float, hipStream_t);
RTM_INSTANTIATE_SNAP(float, float)
RTM_INSTANTIATE_SNAP(float, _Float16)
RTM_INSTANTIATE_SNAP(_Float16, float)
RTM_INSTANTIATE_SNAP(_Float16, _Float16)
RTM_INSTANTIATE_SNAP(__bf16, float)
RTM_INSTANTIATE_SNAP(__bf16, _Float16)

}  // namespace rtmhip
