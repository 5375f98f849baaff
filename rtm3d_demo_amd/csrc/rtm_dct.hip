// rtm_dct.hip -- transform coding of source-wavefield snapshots (no reference
// counterpart: the reference keeps the source field by HBC reconstruction or
// as exact snapshots).
//
// The extended grid is cut into 8 x 8 x 32 bricks (x, y, z; z fastest, the
// z origin at 0, so a brick's z lines are 128-byte rows); a brick's cells
// outside the imaging range z in [z0, z1) (or past the grid) enter as 0.
// Each brick goes through the orthonormal 3D DCT-II (separable: 32-point
// along z, 8-point along y and x).  One threshold per snapshot keeps the
// largest coefficients of the whole snapshot within a budget of K: the
// threshold is a power of two, the lowest whose kept count fits (from
// exponent histograms of the coefficients), so between about K/2 and K
// coefficients survive.  A kept coefficient is one 32-bit entry: its index
// in the brick (11 bits) and its value divided by the brick's largest
// magnitude, in FP16; a brick table holds each brick's first entry, entry
// count and scale.  Decoding scatters a brick's entries into a zero brick,
// runs the inverse transform and either writes the cells or applies the
// imaging condition img += w * (v * PR) directly.
//
// A brick is handled by one block of 256 threads in LDS; blocks loop over
// bricks and add their exponent histograms to one 256-bin histogram.
#include "rtm_internal.hpp"

#include <algorithm>

namespace rtmhip {
namespace {

constexpr int kDctThreads = 256;
constexpr int kLine = 33;                   // padded z line in LDS
constexpr int kTile = 8 * 8 * kLine;        // one brick in LDS
constexpr int kCells = 8 * 8 * 32;          // 2048 cells / coefficients per brick

// The default basis, orthonormal DCT-II matrices D[u][i] = s(u) cos(pi (2i +
// 1) u / 2n), s(0) = sqrt(1/n), s(u > 0) = sqrt(2/n), evaluated at compile
// time (cosine by range reduction and its Taylor series in double).  The
// transforms read the basis through a pointer with wave-uniform indices, so
// its entries come through scalar loads.
constexpr double kPi = 3.14159265358979323846;
constexpr double ccos(double x)
{
    const double two_pi = 2.0 * kPi;
    x -= (double)(long long)(x / two_pi) * two_pi;
    if (x > kPi) x -= two_pi;
    double term = 1.0, sum = 1.0;
    for (int k = 1; k < 40; ++k) {
        term *= -x * x / (double)((2 * k - 1) * (2 * k));
        sum += term;
    }
    return sum;
}
constexpr BrickBasis make_dct_basis()
{
    BrickBasis t{};
    for (int u = 0; u < 32; ++u)
        for (int i = 0; i < 32; ++i)
            t.z[u][i] = (float)((u ? 0.25 : 0.17677669529663688) *
                                ccos(kPi * (double)((2 * i + 1) * u) / 64.0));
    for (int u = 0; u < 8; ++u)
        for (int i = 0; i < 8; ++i) {
            t.x[u][i] = (float)((u ? 0.5 : 0.35355339059327379) *
                                ccos(kPi * (double)((2 * i + 1) * u) / 16.0));
            t.y[u][i] = t.x[u][i];
        }
    return t;
}
__constant__ BrickBasis kDct = make_dct_basis();

// Basis entries: the caller's (L, read through B with wave-uniform indices,
// so they come through scalar loads) or the compile-time DCT-II table, whose
// entries at constant indices the compiler folds into the instructions.
template <bool L>
__device__ __forceinline__ float bz(const BrickBasis *__restrict__ B, int u, int i)
{
    if constexpr (L) return B->z[u][i];
    else return kDct.z[u][i];
}
template <bool L>
__device__ __forceinline__ float bx(const BrickBasis *__restrict__ B, int u, int i)
{
    if constexpr (L) return B->x[u][i];
    else return kDct.x[u][i];
}
template <bool L>
__device__ __forceinline__ float by(const BrickBasis *__restrict__ B, int u, int i)
{
    if constexpr (L) return B->y[u][i];
    else return kDct.y[u][i];
}

__global__ void k_put_basis(BrickBasis *__restrict__ dst)
{
    const int *src = reinterpret_cast<const int *>(&kDct);
    int *d = reinterpret_cast<int *>(dst);
    for (int i = threadIdx.x; i < (int)(sizeof(BrickBasis) / 4); i += blockDim.x) d[i] = src[i];
}

__device__ __forceinline__ void brick_coords(long long br, const DctLayout &L, int &x0, int &y0,
                                             int &z0)
{
    const int bz = (int)(br % L.nbz);
    const long long r = br / L.nbz;
    const int by = (int)(r % L.nby);
    const int bx = (int)(r / L.nby);
    x0 = 8 * bx;
    y0 = 8 * by;
    z0 = 32 * bz;
}

// Forward transform of the brick in t (layout [x][y][z], z padded to 33);
// leaves coefficient (u, v, w) at t[(u*8 + v)*33 + w], except that the x
// pass's outputs stay in the thread's out (x index u = 0..7 at y = tid>>5,
// z = tid&31).  z pass: wave q computes outputs 8q .. 8q+7 of the 64 lines
// (lane = line), so every matrix index is wave-uniform.
template <bool L>
__device__ __forceinline__ void dct3_forward(float *t, const BrickBasis *__restrict__ B,
                                             float (&out)[8])
{
    const int tid = threadIdx.x;
    {   // z
        const int l = tid & 63;
        const int q = __builtin_amdgcn_readfirstlane(tid >> 6);
        float in[32];
#pragma unroll
        for (int k = 0; k < 32; ++k) in[k] = t[l * kLine + k];
        float o[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            float a = 0.0f;
#pragma unroll
            for (int k = 0; k < 32; ++k) a = __builtin_fmaf(bz<L>(B, 8 * q + r, k), in[k], a);
            o[r] = a;
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 8; ++r) t[l * kLine + 8 * q + r] = o[r];
        __syncthreads();
    }
    {   // y: line (i, k)
        const int i = tid >> 5, k = tid & 31;
        float in[8], o[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) in[j] = t[(i * 8 + j) * kLine + k];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            float a = 0.0f;
#pragma unroll
            for (int j = 0; j < 8; ++j) a = __builtin_fmaf(by<L>(B, u, j), in[j], a);
            o[u] = a;
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 8; ++u) t[(i * 8 + u) * kLine + k] = o[u];
        __syncthreads();
    }
    {   // x: line (j, k)
        const int j = tid >> 5, k = tid & 31;
        float in[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) in[i] = t[(i * 8 + j) * kLine + k];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            float a = 0.0f;
#pragma unroll
            for (int i = 0; i < 8; ++i) a = __builtin_fmaf(bx<L>(B, u, i), in[i], a);
            out[u] = a;
        }
    }
}

// Inverse transform in place (coefficients in t as dct3_forward leaves them).
template <bool L>
__device__ __forceinline__ void dct3_inverse(float *t, const BrickBasis *__restrict__ B)
{
    const int tid = threadIdx.x;
    {   // x
        const int j = tid >> 5, k = tid & 31;
        float in[8], o[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) in[u] = t[(u * 8 + j) * kLine + k];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            float a = 0.0f;
#pragma unroll
            for (int u = 0; u < 8; ++u) a = __builtin_fmaf(bx<L>(B, u, i), in[u], a);
            o[i] = a;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 8; ++i) t[(i * 8 + j) * kLine + k] = o[i];
        __syncthreads();
    }
    {   // y
        const int i = tid >> 5, k = tid & 31;
        float in[8], o[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) in[u] = t[(i * 8 + u) * kLine + k];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            float a = 0.0f;
#pragma unroll
            for (int u = 0; u < 8; ++u) a = __builtin_fmaf(by<L>(B, u, j), in[u], a);
            o[j] = a;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 8; ++j) t[(i * 8 + j) * kLine + k] = o[j];
        __syncthreads();
    }
    {   // z
        const int l = tid & 63;
        const int q = __builtin_amdgcn_readfirstlane(tid >> 6);
        float in[32];
#pragma unroll
        for (int w = 0; w < 32; ++w) in[w] = t[l * kLine + w];
        float o[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            float a = 0.0f;
#pragma unroll
            for (int w = 0; w < 32; ++w) a = __builtin_fmaf(bz<L>(B, w, 8 * q + r), in[w], a);
            o[r] = a;
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 8; ++r) t[l * kLine + 8 * q + r] = o[r];
        __syncthreads();
    }
}

// Cell e (0..2047) of a brick: x = e >> 8, y = (e >> 5) & 7, z = e & 31
// (the coefficient index uses the same order).
__device__ __forceinline__ int tile_index(int e) { return (e >> 5) * kLine + (e & 31); }

// Transform of every brick: coefficients to coef[brick][2048] (index
// (u*8 + v)*32 + w); the exponent histogram of the non-zero magnitudes is
// accumulated in LDS per block and added to hist[256] (zeroed beforehand).
template <typename T>
__global__ void __launch_bounds__(kDctThreads)
    k_dct_encode(const T *__restrict__ F, Grid g, int z0, int z1, DctLayout L,
                 const BrickBasis *__restrict__ B, float *__restrict__ coef,
                 unsigned *__restrict__ hist)
{
    __shared__ float t[kTile];
    __shared__ unsigned h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int tid = threadIdx.x;
    const bool learned = B->learned != 0;
    for (long long br = blockIdx.x; br < L.nbr; br += gridDim.x) {
        int bx0, by0, bz0;
        brick_coords(br, L, bx0, by0, bz0);
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            const int e = tid + kDctThreads * m;
            const int x = bx0 + (e >> 8), y = by0 + ((e >> 5) & 7), z = bz0 + (e & 31);
            float v = 0.0f;
            if (x < g.nx && y < g.ny && z >= z0 && z < z1)
                v = wld(F[(long long)x * g.sx + (long long)y * g.sy + z]);
            t[tile_index(e)] = v;
        }
        __syncthreads();
        float out[8];
        if (learned)
            dct3_forward<true>(t, B, out);
        else
            dct3_forward<false>(t, B, out);
        float *cb = coef + br * kCells;
        const int j = tid >> 5, k = tid & 31;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            cb[(u * 8 + j) * 32 + k] = out[u];
            const unsigned bits = __builtin_bit_cast(unsigned, out[u]) & 0x7fffffffu;
            if (bits) atomicAdd(&h[bits >> 23], 1u);
        }
        __syncthreads();  // the next brick reuses t
    }
    if (h[tid]) atomicAdd(&hist[tid], h[tid]);
}

__global__ void k_dct_clear(unsigned *__restrict__ hist) { hist[threadIdx.x] = 0; }

// Threshold of one snapshot: meta[0] = the lowest power-of-two magnitude
// (as float bits) whose kept count (|c| >= it) is at most K, meta[1] = that
// count, meta[2] = 0 (the entry counter of the packing pass).
__global__ void __launch_bounds__(kDctThreads)
    k_dct_select(const unsigned *__restrict__ hist, long long K, unsigned *__restrict__ meta)
{
    __shared__ unsigned long long tot[256];
    tot[threadIdx.x] = hist[threadIdx.x];
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long cum = 0;
        unsigned thr = 0x7f800000u;  // nothing kept below infinity if bin 255 overflows
        int e = 255;
        for (; e >= 0; --e) {
            if (cum + tot[e] > (unsigned long long)K) break;
            cum += tot[e];
        }
        // bins above e are kept; bin 0 (subnormal magnitudes) starts at bits 1
        thr = e < 0 ? 1u : ((unsigned)(e + 1) << 23);
        meta[0] = thr;
        meta[1] = (unsigned)cum;
        meta[2] = 0;
    }
}

// Packing: the coefficients with |c| >= threshold, per brick in index order,
// appended to the entry pool (a brick's range reserved with one atomic), and
// the brick table {first entry, count, scale}.
__global__ void __launch_bounds__(kDctThreads)
    k_dct_pack(const float *__restrict__ coef, DctLayout L, unsigned *__restrict__ meta,
               unsigned *__restrict__ entries, unsigned *__restrict__ table)
{
    __shared__ unsigned wsum[kDctThreads / 64];
    __shared__ float wmax[kDctThreads / 64];
    __shared__ unsigned sbase;
    const unsigned thr = meta[0];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (long long br = blockIdx.x; br < L.nbr; br += gridDim.x) {
        const float *cb = coef + br * kCells + 8 * tid;  // this thread: 8 consecutive
        float c[8];
        unsigned keep = 0, n = 0;
        float mx = 0.0f;
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            c[m] = cb[m];
            const unsigned bits = __builtin_bit_cast(unsigned, c[m]) & 0x7fffffffu;
            if (bits >= thr && bits) {
                keep |= 1u << m;
                ++n;
            }
            mx = fmaxf(mx, fabsf(c[m]));
        }
        // exclusive prefix of n over the block (index order = thread order)
        unsigned incl = n;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const unsigned o = __shfl_up(incl, d, 64);
            if (lane >= d) incl += o;
        }
        float wm = mx;
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) wm = fmaxf(wm, __shfl_xor(wm, d, 64));
        if (lane == 63) wsum[wv] = incl;
        if (lane == 0) wmax[wv] = wm;
        __syncthreads();
        unsigned before = 0, total = 0;
        float scale = 0.0f;
#pragma unroll
        for (int w = 0; w < kDctThreads / 64; ++w) {
            if (w < wv) before += wsum[w];
            total += wsum[w];
            scale = fmaxf(scale, wmax[w]);
        }
        if (tid == 0) {
            const unsigned base = total ? atomicAdd(&meta[2], total) : 0u;
            sbase = base;
            table[br * 3 + 0] = base;
            table[br * 3 + 1] = total;
            table[br * 3 + 2] = __builtin_bit_cast(unsigned, scale);
        }
        __syncthreads();
        unsigned pos = sbase + before + incl - n;
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            if ((keep >> m) & 1u) {
                // a division, not a product with 1/scale: the reciprocal of a
                // subnormal scale overflows (FP32 fields reach subnormal
                // magnitudes far from the source), the quotient stays <= 1
                const _Float16 hv = (_Float16)(c[m] / scale);
                entries[pos++] = (unsigned)(8 * tid + m) |
                                 ((unsigned)__builtin_bit_cast(unsigned short, hv) << 16);
            }
        }
        __syncthreads();  // wsum / sbase reused by the next brick
    }
}

// Decoding: each brick rebuilt from its entries; IMG: img += w * (v * PR)
// on the imaging range, else out = v there (FP32, the grid's layout).
template <typename T, bool IMG>
__global__ void __launch_bounds__(kDctThreads)
    k_dct_decode(const unsigned *__restrict__ entries, const unsigned *__restrict__ table,
                 DctLayout L, const BrickBasis *__restrict__ B, Grid g, int z0, int z1,
                 const T *__restrict__ PR, float *__restrict__ img, float w,
                 float *__restrict__ out)
{
    __shared__ float t[kTile];
    const int tid = threadIdx.x;
    const bool learned = B->learned != 0;
    for (long long br = blockIdx.x; br < L.nbr; br += gridDim.x) {
        for (int e = tid; e < kTile; e += kDctThreads) t[e] = 0.0f;
        __syncthreads();
        const unsigned base = table[br * 3 + 0], n = table[br * 3 + 1];
        const float scale = __builtin_bit_cast(float, table[br * 3 + 2]);
        for (unsigned e = tid; e < n; e += kDctThreads) {
            const unsigned u = entries[base + e];
            const float v = (float)__builtin_bit_cast(_Float16, (unsigned short)(u >> 16));
            t[tile_index((int)(u & 0x7ffu))] = v * scale;
        }
        // the imaging operands of this thread's cells, loaded before the
        // inverse transform so that their latency overlaps it
        int bx0, by0, bz0;
        brick_coords(br, L, bx0, by0, bz0);
        long long o[8];
        bool in[8];
        // PR as raw bits in 32-bit registers (16-bit values: zero-extending
        // loads, not half-register ones), widened after the transform
        unsigned pr[8];
        float im[8];
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            const int e = tid + kDctThreads * m;
            const int x = bx0 + (e >> 8), y = by0 + ((e >> 5) & 7), z = bz0 + (e & 31);
            in[m] = x < g.nx && y < g.ny && z >= z0 && z < z1;
            o[m] = (long long)x * g.sx + (long long)y * g.sy + z;
            pr[m] = 0u;
            im[m] = 0.0f;
            if (IMG && in[m]) {
                if constexpr (sizeof(T) == 2)
                    pr[m] = reinterpret_cast<const unsigned short *>(PR)[o[m]];
                else
                    pr[m] = reinterpret_cast<const unsigned *>(PR)[o[m]];
                im[m] = img[o[m]];
            }
        }
        __syncthreads();
        if (learned)
            dct3_inverse<true>(t, B);
        else
            dct3_inverse<false>(t, B);
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            if (in[m]) {
                const float v = t[tile_index(tid + kDctThreads * m)];
                if (IMG) {
                    float prv;
                    if constexpr (sizeof(T) == 2)
                        prv = wld(__builtin_bit_cast(T, (unsigned short)pr[m]));
                    else
                        prv = __builtin_bit_cast(float, pr[m]);
                    const float p = v * prv;
                    img[o[m]] = im[m] + w * p;
                } else {
                    out[o[m]] = v;
                }
            }
        }
        __syncthreads();  // the next brick clears t
    }
}

// blocks per launch: each loops over bricks (up to 16 per CU on 256 CUs)
inline unsigned dct_blocks(long long nbr) { return (unsigned)std::min<long long>(nbr, 4096); }

}  // namespace

hipError_t dct_default_basis(BrickBasis *dst, hipStream_t s)
{
    hipLaunchKernelGGL(k_put_basis, dim3(1), dim3(256), 0, s, dst);
    return hipGetLastError();
}

DctLayout dct_layout(const Grid &g)
{
    DctLayout L{};
    L.nbx = (g.nx + 7) / 8;
    L.nby = (g.ny + 7) / 8;
    L.nbz = (g.nz + 31) / 32;
    L.nbr = (long long)L.nbx * L.nby * L.nbz;
    L.nblk = (int)dct_blocks(L.nbr);
    return L;
}

template <typename T>
hipError_t launch_dct_encode(const T *F, const Grid &g, int z0, int z1, const DctLayout &L,
                             const BrickBasis *B, long long K, float *coef, unsigned *hist,
                             unsigned *meta, unsigned *entries, unsigned *table, hipStream_t s)
{
    if (L.nbr <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dct_clear, dim3(1), dim3(256), 0, s, hist);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_dct_encode<T>, dim3(L.nblk), dim3(kDctThreads), 0, s, F, g, z0, z1, L, B,
                       coef, hist);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_dct_select, dim3(1), dim3(kDctThreads), 0, s, hist, K, meta);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_dct_pack, dim3(L.nblk), dim3(kDctThreads), 0, s, coef, L, meta, entries,
                       table);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_dct_decode(const unsigned *entries, const unsigned *table, const DctLayout &L,
                             const BrickBasis *B, const Grid &g, int z0, int z1, const T *PR,
                             float *img, float w, float *out, hipStream_t s)
{
    if (L.nbr <= 0) return hipSuccess;
    if (PR)
        hipLaunchKernelGGL((k_dct_decode<T, true>), dim3(L.nblk), dim3(kDctThreads), 0, s,
                           entries, table, L, B, g, z0, z1, PR, img, w, out);
    else
        hipLaunchKernelGGL((k_dct_decode<T, false>), dim3(L.nblk), dim3(kDctThreads), 0, s,
                           entries, table, L, B, g, z0, z1, PR, img, w, out);
    return hipGetLastError();
}

#define RTM_INSTANTIATE_DCT(T)                                                                 \
    template hipError_t launch_dct_encode<T>(const T *, const Grid &, int, int,                \
                                             const DctLayout &, const BrickBasis *, long long,   \
                                             float *, unsigned *, unsigned *, unsigned *,        \
                                             unsigned *, hipStream_t);                           \
    template hipError_t launch_dct_decode<T>(const unsigned *, const unsigned *,               \
                                             const DctLayout &, const BrickBasis *,              \
                                             const Grid &, int, int, const T *, float *, float,  \
                                             float *, hipStream_t);
RTM_INSTANTIATE_DCT(float)
RTM_INSTANTIATE_DCT(_Float16)
RTM_INSTANTIATE_DCT(__bf16)

}  // namespace rtmhip
