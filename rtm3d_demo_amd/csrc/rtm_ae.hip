// rtm_ae.hip -- kernels of the snapshot autoencoder (learned compression of
// source-wavefield snapshots; no reference counterpart).
//
// A snapshot region is cut into 32^3 bricks, each normalised by its largest
// magnitude and stored as FP16 in "AE order": element e of a brick is cell
// (8 ax + 2 mx + lx, 8 ay + 2 my + ly, 8 az + 2 mz + lz) with
//   e = ((a * 64 + m) * 8 + l),  a = (ax*4+ay)*4+az,  m = (mx*4+my)*4+mz,
//   l = (lx*2+ly)*2+lz.
// In this order every layer of the network is a plain row-major GEMM over
// non-overlapping patches:
//   L1  2^3 patches, 1 -> 8 channels, tanh    [4096 rows x 8]  * [8 x 8]
//   L2  4^3 patches, 8 -> 32 channels (code)   [64 rows x 512]  * [512 x 32]
//   L3  transposed L2, 32 -> 8, tanh           [64 rows x 32]   * [32 x 512]
//   L4  transposed L1, 8 -> 1                  [4096 rows x 8]  * [8 x 8]
// so one brick of 32768 FP16 values (64 KB) becomes 2048 FP16 codes (4 KB)
// plus its scale.  The GEMMs run on the matrix cores (MFMA 16x16x32, FP16
// operands, FP32 accumulation) for inference and for training (forward,
// backward, split-K weight gradients).
#include "rtm_internal.hpp"

#include <algorithm>

namespace rtmhip {
namespace {

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));

constexpr int kGemmThreads = 256;  // 4 waves; 64 x 64 tile of C, 32 x 32 per wave
constexpr int kBrick = 32768;

__device__ __forceinline__ float load_elem(const void *p, long long i, int is_half)
{
    return is_half ? (float)static_cast<const _Float16 *>(p)[i] : static_cast<const float *>(p)[i];
}

// C(m, n) (+)= sum_k A(m, k) B(k, n), A(m, k) = A[m*sam + k*sak] and
// B(k, n) = B[k*sbk + n*sbn] (either FP32 or FP16 in memory, FP16 operands).
// Epilogue: + bias[n], tanh (act 1); or atomic accumulation of a split-K
// partial (accumulate != 0; blockIdx.z selects the K range).
//
// gfx950's MFMA 16x16x32 (FP16 in, FP32 accumulate): lane l supplies row
// l % 16 of A and column l % 16 of B with eight k values each, k-slot group
// l / 16; A and B use the same k assignment, so the sum over k is the same
// whichever k each slot of the instruction stands for.  D(i, j) of a tile is
// lane j + 16 * (i / 4), item i % 4.  Both operands sit in LDS k-contiguous
// (B transposed), so each lane reads its eight values as one 16-byte load.
__global__ void __launch_bounds__(kGemmThreads) k_gemm(AeGemm p)
{
    constexpr int KT = 32, KP = KT + 8;  // k per tile; padded LDS row (80 B)
    __shared__ __align__(16) _Float16 As[64][KP];
    __shared__ __align__(16) _Float16 Bt[64][KP];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const long long m0 = (long long)blockIdx.x * 64;
    const int n0 = blockIdx.y * 64;
    const long long kb = (long long)blockIdx.z * p.kchunk;
    const long long ke = min((long long)p.K, kb + p.kchunk);
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    float4_t acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = float4_t{0.f, 0.f, 0.f, 0.f};
    for (long long k0 = kb; k0 < ke; k0 += KT) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int idx = tid * 8 + q;  // 0 .. 2047
            {
                const int mm = idx >> 5, kk = idx & 31;
                const long long m = m0 + mm, k = k0 + kk;
                float v = 0.f;
                if (m < p.M && k < ke) v = load_elem(p.A, m * p.sam + k * p.sak, p.a_half);
                As[mm][kk] = (_Float16)v;
            }
            {
                const int kk = idx >> 6, nn = idx & 63;
                const long long k = k0 + kk;
                const int n = n0 + nn;
                float v = 0.f;
                if (n < p.N && k < ke) v = load_elem(p.B, k * p.sbk + (long long)n * p.sbn, p.b_half);
                Bt[nn][kk] = (_Float16)v;
            }
        }
        __syncthreads();
        const int r = lane & 15, kq = (lane >> 4) * 8;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const half8_t a = *reinterpret_cast<const half8_t *>(&As[wm + 16 * i + r][kq]);
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const half8_t b = *reinterpret_cast<const half8_t *>(&Bt[wn + 16 * j + r][kq]);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, acc[i][j], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const long long m = m0 + wm + 16 * i + 4 * (lane >> 4) + v;
                const int n = n0 + wn + 16 * j + (lane & 15);
                if (m >= p.M || n >= p.N) continue;
                float r = acc[i][j][v];
                float *c = p.C + m * p.ldc + n;
                if (p.accumulate) {
                    atomicAdd(c, r);
                } else {
                    if (p.bias) r += p.bias[n];
                    if (p.act == 1) r = tanhf(r);
                    *c = r;
                }
            }
}

// dY = Y - X (the gradient of 0.5 * sum (Y - X)^2), loss += sum (Y - X)^2.
__global__ void __launch_bounds__(256)
    k_ae_residual(const float *__restrict__ Y, const _Float16 *__restrict__ X,
                  float *__restrict__ dY, long long n, float *__restrict__ loss)
{
    __shared__ float red[256];
    float s = 0.f;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (long long)gridDim.x * 256) {
        const float d = Y[i] - (float)X[i];
        dY[i] = d;
        s += d * d;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}

// dA = dH * (1 - H^2): back through tanh, in place.
__global__ void __launch_bounds__(256)
    k_tanh_back(float *__restrict__ dH, const float *__restrict__ H, long long n)
{
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (long long)gridDim.x * 256)
        dH[i] = dH[i] * (1.0f - H[i] * H[i]);
}

// db[c] += sum over rows of D[row][c] (rows x cols, row-major; cols divides
// 256 or is a multiple of 256).  A block walks its row range in row-major
// order, so every load is coalesced; thread t keeps the partial sums of the
// columns t, t + 256, ... (or of column t % cols), then one block reduction.
__global__ void __launch_bounds__(256)
    k_colsum(const float *__restrict__ D, long long rows, int cols, float *__restrict__ db,
             long long rows_per_block)
{
    __shared__ float red[256];
    const long long r0 = (long long)blockIdx.x * rows_per_block;
    const long long r1 = min(rows, r0 + rows_per_block);
    const long long e0 = r0 * cols, e1 = r1 * cols;
    const int t = threadIdx.x;
    if (cols <= 256) {
        float acc = 0.f;
        for (long long e = e0 + t; e < e1; e += 256) acc += D[e];
        red[t] = acc;
        __syncthreads();
        if (t < cols) {
            float s = 0.f;
            for (int k = t; k < 256; k += cols) s += red[k];
            atomicAdd(&db[t], s);
        }
    } else {
        for (int c = t; c < cols; c += 256) {
            float acc = 0.f;
            for (long long e = e0 + c; e < e1; e += cols) acc += D[e];
            atomicAdd(&db[c], acc);
        }
    }
}

// Adam (Kingma & Ba 2015) with bias correction folded into the step size.
__global__ void __launch_bounds__(256)
    k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
           float *__restrict__ v, int n, float lr_t, float b1, float b2, float eps)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] = w[i] - lr_t * mi / (sqrtf(vi) + eps);
}

__global__ void __launch_bounds__(256)
    k_gather_bricks(const _Float16 *__restrict__ src, const int *__restrict__ idx,
                    _Float16 *__restrict__ dst, int n)
{
    const long long total = (long long)n * kBrick;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < total;
         i += (long long)gridDim.x * 256) {
        const long long b = i / kBrick, e = i % kBrick;
        dst[i] = src[(long long)idx[b] * kBrick + e];
    }
}

__global__ void __launch_bounds__(256)
    k_f32_to_f16(const float *__restrict__ a, _Float16 *__restrict__ b, long long n)
{
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (long long)gridDim.x * 256)
        b[i] = (_Float16)a[i];
}

__global__ void __launch_bounds__(256)
    k_f16_to_f32(const _Float16 *__restrict__ a, float *__restrict__ b, long long n)
{
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (long long)gridDim.x * 256)
        b[i] = (float)a[i];
}

// Cell of element e of a brick (AE order).
__device__ __forceinline__ void brick_cell(int e, int &x, int &y, int &z)
{
    const int l = e & 7, mi = (e >> 3) & 63, ai = e >> 9;
    x = 8 * (ai >> 4) + 2 * (mi >> 4) + (l >> 2);
    y = 8 * ((ai >> 2) & 3) + 2 * ((mi >> 2) & 3) + ((l >> 1) & 1);
    z = 8 * (ai & 3) + 2 * (mi & 3) + (l & 1);
}

// One block per brick: scale = max |v| over the brick's in-grid cells, then
// the FP16 values v * (1 / scale) (zero outside the grid, and the whole
// brick zero with scale 0 when max |v| is zero or subnormal).
__global__ void __launch_bounds__(256)
    k_bricks(const float *__restrict__ F, Grid g, AeRegion r, _Float16 *__restrict__ out,
             float *__restrict__ scales)
{
    __shared__ float red[256];
    const int b = blockIdx.x;
    const int bz = b % r.nbz, by = (b / r.nbz) % r.nby, bx = b / (r.nbz * r.nby);
    const int gx0 = r.x0 + 32 * bx, gy0 = r.y0 + 32 * by, gz0 = r.z0 + 32 * bz;
    float mx = 0.f;
    for (int e = threadIdx.x; e < kBrick; e += 256) {
        int x, y, z;
        brick_cell(e, x, y, z);
        x += gx0;
        y += gy0;
        z += gz0;
        if (x < r.x1 && y < r.y1 && z < r.z1)
            mx = fmaxf(mx, fabsf(F[(long long)x * g.sx + (long long)y * g.sy + z]));
    }
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + h]);
        __syncthreads();
    }
    // a brick whose largest magnitude is below the smallest normal float
    // (1/s would overflow) is stored as zero
    const float s = red[0] >= 1.17549435e-38f ? red[0] : 0.f;
    if (threadIdx.x == 0) scales[b] = s;
    const float inv = s > 0.f ? 1.0f / s : 0.f;
    _Float16 *o = out + (long long)b * kBrick;
    for (int e = threadIdx.x; e < kBrick; e += 256) {
        int x, y, z;
        brick_cell(e, x, y, z);
        x += gx0;
        y += gy0;
        z += gz0;
        float v = 0.f;  // +0 outside the grid and in a zero brick
        if (s > 0.f && x < r.x1 && y < r.y1 && z < r.z1)
            v = F[(long long)x * g.sx + (long long)y * g.sy + z] * inv;
        o[e] = wst<_Float16>(v);  // the FP32 product rounded, then narrowed
    }
}

// The inverse: F(cell) = Y[brick][e] * scale for every in-grid cell of the
// region (Y in AE order, FP32).
__global__ void __launch_bounds__(256)
    k_unbricks(const float *__restrict__ Y, const float *__restrict__ scales, Grid g, AeRegion r,
               float *__restrict__ F, int b0, int nb)
{
    const long long total = (long long)nb * kBrick;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < total;
         i += (long long)gridDim.x * 256) {
        const int bl = (int)(i / kBrick), e = (int)(i % kBrick);
        const int b = b0 + bl;
        const int bz = b % r.nbz, by = (b / r.nbz) % r.nby, bx = b / (r.nbz * r.nby);
        int x, y, z;
        brick_cell(e, x, y, z);
        x += r.x0 + 32 * bx;
        y += r.y0 + 32 * by;
        z += r.z0 + 32 * bz;
        if (x < r.x1 && y < r.y1 && z < r.z1)
            F[(long long)x * g.sx + (long long)y * g.sy + z] = Y[i] * scales[b];
    }
}

inline unsigned grid_for(long long n)
{
    long long b = (n + 255) / 256;
    return (unsigned)(b > 16384 ? 16384 : (b ? b : 1));
}

}  // namespace

hipError_t launch_ae_gemm(const AeGemm &p, hipStream_t s)
{
    if (p.M <= 0 || p.N <= 0 || p.K <= 0) return hipSuccess;
    AeGemm q = p;
    long long splits = 1;
    if (q.accumulate) {
        if (q.kchunk <= 0) {
            // split K so that the launch has about 1024 blocks, 256..65536 per split
            const long long tiles = ((q.M + 63) / 64) * ((q.N + 63) / 64);
            const long long want = std::max(1LL, 1024 / std::max(1LL, tiles));
            q.kchunk = (int)std::min(65536LL, std::max(256LL, (q.K + want - 1) / want));
        }
        q.kchunk = (q.kchunk + 31) / 32 * 32;
        splits = (q.K + q.kchunk - 1) / q.kchunk;
    } else {
        q.kchunk = (int)((q.K + 31) / 32 * 32);
    }
    const long long gm = (q.M + 63) / 64, gn = (q.N + 63) / 64;
    if (gm > 0x7fffffffLL || gn > 65535 || splits > 65535) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_gemm, dim3((unsigned)gm, (unsigned)gn, (unsigned)splits),
                       dim3(kGemmThreads), 0, s, q);
    return hipGetLastError();
}

hipError_t launch_ae_residual(const float *Y, const void *X, float *dY, long long n, float *loss,
                              hipStream_t s)
{
    // at most 1024 blocks: one atomic per block on the single loss word
    hipLaunchKernelGGL(k_ae_residual, dim3(std::min(1024u, grid_for(n))), dim3(256), 0, s, Y,
                       static_cast<const _Float16 *>(X), dY, n, loss);
    return hipGetLastError();
}

hipError_t launch_tanh_back(float *dH, const float *H, long long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_tanh_back, dim3(grid_for(n)), dim3(256), 0, s, dH, H, n);
    return hipGetLastError();
}

hipError_t launch_colsum(const float *D, long long rows, int cols, float *db, hipStream_t s)
{
    if (cols <= 0 || rows <= 0) return hipSuccess;
    if (!(cols <= 256 ? 256 % cols == 0 : cols % 256 == 0)) return hipErrorInvalidValue;
    // about 16K elements per block
    const long long rpb = std::max(1LL, 16384LL / cols);
    const long long nb = (rows + rpb - 1) / rpb;
    if (nb > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_colsum, dim3((unsigned)nb), dim3(256), 0, s, D, rows, cols, db, rpb);
    return hipGetLastError();
}

hipError_t launch_adam(float *w, const float *g, float *m, float *v, int n, float lr_t, float b1,
                       float b2, float eps, hipStream_t s)
{
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, w, g, m, v, n,
                       lr_t, b1, b2, eps);
    return hipGetLastError();
}

hipError_t launch_gather_bricks(const void *src, const int *idx, void *dst, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_gather_bricks, dim3(grid_for((long long)n * kBrick)), dim3(256), 0, s,
                       static_cast<const _Float16 *>(src), idx, static_cast<_Float16 *>(dst), n);
    return hipGetLastError();
}

hipError_t launch_f32_to_f16(const float *a, void *b, long long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_f32_to_f16, dim3(grid_for(n)), dim3(256), 0, s, a,
                       static_cast<_Float16 *>(b), n);
    return hipGetLastError();
}

hipError_t launch_f16_to_f32(const void *a, float *b, long long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_f16_to_f32, dim3(grid_for(n)), dim3(256), 0, s,
                       static_cast<const _Float16 *>(a), b, n);
    return hipGetLastError();
}

hipError_t launch_bricks(const float *F, const Grid &g, const AeRegion &r, void *out,
                         float *scales, hipStream_t s)
{
    const long long nb = (long long)r.nbx * r.nby * r.nbz;
    if (nb <= 0) return hipSuccess;
    if (nb > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_bricks, dim3((unsigned)nb), dim3(256), 0, s, F, g, r,
                       static_cast<_Float16 *>(out), scales);
    return hipGetLastError();
}

hipError_t launch_unbricks(const float *Y, const float *scales, const Grid &g, const AeRegion &r,
                           float *F, int b0, int nb, hipStream_t s)
{
    if (nb <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_unbricks, dim3(grid_for((long long)nb * kBrick)), dim3(256), 0, s, Y,
                       scales, g, r, F, b0, nb);
    return hipGetLastError();
}

}  // namespace rtmhip
