// rtm_internal.hpp -- shared declarations of the gfx950 kernels and their
// launchers.  Not part of the public C ABI (include/rtm_hip.h).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <type_traits>

namespace rtmhip {

// Wavefield storage formats.  Arithmetic is always FP32: a wavefield value
// is widened on load and rounded (nearest-even) on store; "rnd" is the
// round trip, used where the reference's operator sequence would have
// stored the value before reading it back.
enum WaveFormat { WAVE_F32 = 0, WAVE_F16 = 1, WAVE_BF16 = 2 };

__device__ __forceinline__ float wld(float v) { return v; }
__device__ __forceinline__ float wld(_Float16 v) { return (float)v; }
__device__ __forceinline__ float wld(__bf16 v) { return (float)v; }
template <typename T>
__device__ __forceinline__ T wst(float v)
{
    // The FP32 value is made opaque first: otherwise the compiler fuses the
    // multiply or add that produced it with the narrowing into one
    // v_fma_mixlo_f16, which rounds the exact result once to FP16 instead of
    // rounding to FP32 and then to FP16 as the operator sequence does.
    if constexpr (!std::is_same<T, float>::value) asm("" : "+v"(v));
    return (T)v;
}
template <typename T>
__device__ __forceinline__ float wrnd(float v)
{
    if constexpr (std::is_same<T, float>::value) return v;
    else return wld(wst<T>(v));
}

// Stencil plane [dim][tap] (rtm-3d/include/RTMGrid.hpp:400-446), passed by
// value so every tap lives in a scalar register.
struct Coefs {
    float c[3][9];
};

// Extended grid with explicit strides: offset = x*sx + y*sy + z.
struct Grid {
    int nx, ny, nz;
    long long sx;  // plane stride (elements)
    int sy;        // row stride (elements)
};

inline Grid dense_grid(int nx, int ny, int nz)
{
    return Grid{nx, ny, nz, (long long)ny * nz, nz};
}

// Epilogue of the fused receiver step: img += ps*pp on z in [z0, z1) except
// at receiver cells, which the receiver kernel handles after injection.
struct ImageFuse {
    float *img;
    const void *ps;  // source wavefield, of the step's storage type
    int z0, z1;
    int rz;             // receiver plane (extended z)
    int rx0, rdx, rnx;  // receiver x = rx0 + i*rdx, i < rnx (extended)
    int ry0, rdy, rny;
};

// Upper-border taper fused into the stencil (HBC steps): P values with
// z < b are tapered once on load, the centre PP value pp_taps times.
struct TaperFuse {
    const float *t;  // device table g[b]
    int b;
    int pp_taps;     // 1 or 2 (all 2: 0 or 1)
    int all;         // 0: the upper taper (taperUpper3D) on load; 1: every border
                     // (taper3D) on load; 2: every border on the step's output (P stored
                     // tapered); 3: the upper taper on the step's output
    int hold;        // all 3: planes z in [b - hold, b) stored untapered (the upper
                     // border history is saved from them, then they are tapered)
};

// Tile configuration of the stencil kernel.
struct StencilTiling {
    int tz, by, ry;  // lanes along z, thread rows, rows per thread
    int xchunk;      // planes marched per block (0 = automatic)
};

// One shot of a batch (many shots advanced in lockstep by single launches),
// as the batched kernels read it from device memory.  The wavefield pairs
// are fixed buffers; which one is "P" and which "PP" at a step is a
// batch-wide parity passed to each launch.
struct BatchShot {
    void *fs[2];         // source wavefield pair (storage type T)
    void *fr[2];         // receiver wavefield pair (backward pass)
    const float *V;      // v^2 dt^2
    float *img;
    void *upb;           // upper-border history [nt][nx][ny][hf], storage type T
    const float *rcv;    // receiver record [rnx][rny][nt]
    const float *ricker; // source samples [nt]
    Grid g;
    long long src_off;   // source cell
    int rx0, ry0, rz, rdx, rdy, rnx, rny;  // receivers (extended coordinates)
    int nzt, nyt, xchunk;                   // waves of the batched stencil
    long long nwaves;                       // nzt * nyt * (x chunks)
    long long blk0;      // first block of this shot in a batched stencil launch
    int fwd0, nfwd;      // first block and work items of this shot in the forward tail
    int bwd0, nbwd;      // ... in the backward tail
};

// Tile shapes compiled into the stencil kernel (tz lanes x by rows, ry rows
// per thread); the default is 64x4x4.
bool stencil_tiling_supported(int tz, int by, int ry);

// Launchers: enqueue on `s`, return the launch status.  T is the wavefield
// storage type (float, _Float16 or __bf16); V, tapers, traces and the image
// are always float.  Instantiated for the three storage types.
template <typename T>
hipError_t launch_stencil(int radius, const T *P, T *PP, const float *V, const Coefs &cf,
                          const Grid &g, const ImageFuse *fuse, const TaperFuse *taper,
                          int *flag, hipStream_t s, const StencilTiling *tiling);
// Backward step of the source and the receiver field in one pass (FP32,
// radius 4): source plain step into PPs, receiver step (tapered when taper
// is given) into PPr, img += Ps_new * PPr_new off the receiver cells.
bool dual_step_supported(int radius, int wave_format);
hipError_t launch_stencil_dual(int radius, const float *Ps, float *PPs, const float *Pr,
                               float *PPr, const float *V, const Coefs &cf, const Grid &g,
                               const ImageFuse &fuse, const TaperFuse *taper, int *flag,
                               hipStream_t s, const StencilTiling *tiling);
template <typename T>
hipError_t launch_taper(T *P, const float *taper, const Grid &g, int blen, bool upper_only,
                        hipStream_t s);
// Snapshots over the imaging range z in [b, nz-b) ([nx][ny][nz-2b], element
// type S: float or _Float16), and the imaging condition from one:
// img += w * (snap * PR) on that range.
template <typename T, typename S>
hipError_t launch_snap_store(const T *P, S *snap, const Grid &g, int b, hipStream_t s);
template <typename T, typename S>
hipError_t launch_snap_image(float *img, const S *snap, const T *PR, const Grid &g, int b,
                             float w, hipStream_t s);
// Transform-coded snapshots (rtm_dct.hip): 8 x 8 x 32 bricks of the grid,
// 3D DCT-II, one power-of-two threshold per snapshot keeping at most K
// coefficients, entries {index in brick (11 bits), FP16 value / brick scale}
// and a brick table {first entry, count, scale bits} (3 x 32 bits).
// Separable orthonormal basis of a brick: rows are the basis vectors along
// z (32), x (8) and y (8); forward coefficient (u, v, w) = sum over the
// brick of x[u][i] y[v][j] z[w][k] f(i, j, k).  dct_default_basis writes the
// DCT-II into device memory.
struct BrickBasis {
    float z[32][32];
    float x[8][8];
    float y[8][8];
    int learned;  // 0: the DCT-II (the kernels use their compile-time copy)
};
hipError_t dct_default_basis(BrickBasis *dst, hipStream_t s);
struct DctLayout {
    int nbx, nby, nbz;  // bricks along x, y, z
    int nblk;           // blocks per launch (each loops over bricks)
    long long nbr;      // bricks
};
DctLayout dct_layout(const Grid &g);
// coef: [nbr][2048] FP32 scratch; hist: 256 words of scratch; meta: 3 words
// (threshold bits, kept count, entry counter); the cells z in [z0, z1) enter.
template <typename T>
hipError_t launch_dct_encode(const T *F, const Grid &g, int z0, int z1, const DctLayout &L,
                             const BrickBasis *B, long long K, float *coef, unsigned *hist,
                             unsigned *meta, unsigned *entries, unsigned *table, hipStream_t s);
// PR != NULL: img += w * (v * PR) on z in [z0, z1); else out = v there.
template <typename T>
hipError_t launch_dct_decode(const unsigned *entries, const unsigned *table, const DctLayout &L,
                             const BrickBasis *B, const Grid &g, int z0, int z1, const T *PR,
                             float *img, float w, float *out, hipStream_t s);
// taperUpper3D of two grids of the same shape in one pass (Q may be NULL).
template <typename T>
hipError_t launch_taper_upper2(T *P, T *Q, const float *taper, const Grid &g, int b,
                               hipStream_t s);
// taper3D of two grids of the same shape in one pass (Q may be NULL: one grid).
template <typename T>
hipError_t launch_taper_all2(T *P, T *Q, const float *taper, const Grid &g, int b,
                             hipStream_t s);
// Upper taper (applied `taps` times) of the slab cells the stencil never
// updates (z < R or the x / y frame of width R), in place.
// With hold > 0 the planes z in [b - hold, b) get one application fewer.
template <typename T>
hipError_t launch_taper_frame(T *P, const float *taper, const Grid &g, int b, int R, int taps,
                              hipStream_t s, int hold = 0);
// taper3D's counterpart of launch_taper_frame (fused every-border taper):
// the cells within R of any grid face, `taps` times in place, the front-right
// edge's cells from the cell (b, y, z) as taper3D reads it.  Runs before the
// step that reads the field's plane b as the step's PP.
template <typename T>
hipError_t launch_taper_frame_all(T *P, const float *taper, const Grid &g, int b, int R, int taps,
                                  hipStream_t s);
// taper3D on a step's output (TaperFuse all 2): the cells with x or y within
// R of a face, `taps` times in place, front-right edge cells excepted
// (launch_taper_frame_xy, before the step); then the front-right edge cells
// from the stored (b, y, z) (launch_taper_quirk_out, after the step).
template <typename T>
hipError_t launch_taper_frame_xy(T *P, const float *taper, const Grid &g, int b, int R, int taps,
                                 hipStream_t s);
template <typename T>
hipError_t launch_taper_quirk_out(T *P, const float *taper, const Grid &g, int b, hipStream_t s);
template <typename T>
hipError_t launch_source(T *PP, long long offset, const float *table, int it, float value,
                         hipStream_t s);
template <typename T>
hipError_t launch_receivers(T *PP, float *rcv, const Grid &g, int rx0, int ry0, int rz, int rdx,
                            int rdy, int rnx, int rny, int ntstep, int lt, bool restore,
                            hipStream_t s);
template <typename T>
hipError_t launch_receivers_image(T *PP, const float *rcv, const Grid &g, int rx0, int ry0,
                                  int rz, int rdx, int rdy, int rnx, int rny, int ntstep, int lt,
                                  float *img, const T *ps, int z0, int z1, int R, hipStream_t s);
template <typename T>
hipError_t launch_upb(T *PP, T *upb, const Grid &g, int blen, int hf, long long lt, bool restore,
                      hipStream_t s);
// The upper-border history saved from PP's planes [blen-hf, blen), then the
// upper taper applied to those planes in place (a step that stores the upper
// taper of its output left them untapered).
template <typename T>
hipError_t launch_upb_save_taper(T *PP, T *upb, const float *taper, const Grid &g, int blen,
                                 int hf, long long lt, hipStream_t s);
template <typename T>
hipError_t launch_image(float *img, const T *PS, const T *PR, const Grid &g, int z0, int z1,
                        hipStream_t s);
hipError_t launch_image_finish(const float *img, float *out, const Grid &g, int blen,
                               const Coefs &fc, hipStream_t s);
// Batched launches over `n` shots (device table `shots`, host copy `hs`
// for the launch geometry).  Stencil: every shot's step in one launch of
// the barrier-free wave kernel with ry rows per wave (6 or 8);
// IMG: receiver step fused with imaging (field fr, source value fs[pspar]),
// else the source field fs.  par selects P = f[par], PP = f[par ^ 1].
template <typename T>
hipError_t launch_stencil_batch(int ry, const BatchShot *shots, const BatchShot *hs, int n,
                                long long nblocks, bool img, int par, int pspar, const Coefs &cf,
                                const TaperFuse *taper, int *flag, hipStream_t s);
// Block count and per-shot wave geometry (nzt, nyt, xchunk, nwaves, blk0) of
// a batched stencil launch, written into hs.
long long plan_stencil_batch(int ry, BatchShot *hs, int n, int R, int blen);
// Forward tail of one HBC step on every shot: frame taper (taps times) of
// PP = fs[par ^ 1], source injection, UPB save at time it; with out (a step
// that stored the upper taper of its output outside the UPB layer) the
// layer is saved and then tapered once, its frame cells tapered taps - 1
// times before the save.
// Backward tail: UPB restore into fs[spar ^ 1], frame taper of
// fr[rpar ^ 1], receiver injection and the imaging condition off the
// stencil's cells.  it = (itp ? *itp : 0) + itoff.
template <typename T>
hipError_t launch_tail_fwd(const BatchShot *shots, int n, int blocks, int par,
                           const float *taper, int b, int R, int hf, int taps, const int *itp,
                           int itoff, hipStream_t s, bool out = false);
template <typename T>
hipError_t launch_tail_bwd(const BatchShot *shots, int n, int blocks, int spar, int rpar,
                           const float *taper, int b, int R, int hf, int taps, int nt,
                           const int *itp, int itoff, hipStream_t s);
// Work items of the tails for one shot, and blocks for a count of items.
long long tail_fwd_items(const Grid &g, int b, int R, int hf);
long long tail_bwd_items(const Grid &g, int b, int R, int hf, int rnx, int rny);
int tail_blocks(long long items);
// Brackets a stream capture: while on, stencil launchers use only launch
// shapes already seen eagerly (no occupancy or attribute call inside the
// capture) and fail with hipErrorStreamCaptureUnsupported on a new one.
void capture_guard(bool on);
// *p += d, or *p = d with set (one thread): the device-resident time index
// of a captured step sequence.
hipError_t launch_tick(int *p, int d, bool set, hipStream_t s);
// Frequency-domain migration: contributions of the frequencies iw0 ..
// iw0+nwstep-1 (< nw) at time index st and the imaging condition, FP32.
hipError_t launch_freq_image(float *img, const float *PS, const float *PR, float *sre,
                             float *sim, float *rre, float *rim, const float *kre,
                             const float *kim, const float *w2, const Grid &g, int b, int nwstep,
                             int iw0, int nw, int nt, int st, hipStream_t s);
hipError_t launch_fill(float *p, size_t n, float v, hipStream_t s);
hipError_t launch_fill_u16(uint16_t *p, size_t n, uint16_t bits, hipStream_t s);
hipError_t launch_stack(float *dst, const float *src, size_t n, hipStream_t s);

// ---- snapshot autoencoder (rtm_ae.hip) ----
// One GEMM of the autoencoder on the matrix cores: C(m, n) (+)= sum_k
// A(m, k) B(k, n) with A(m, k) = A[m*sam + k*sak], B(k, n) = B[k*sbk +
// n*sbn] (FP32 or FP16 in memory, FP16 operands, FP32 accumulation).
struct AeGemm {
    const void *A;
    long long sam, sak;
    int a_half;
    const void *B;
    long long sbk, sbn;
    int b_half;
    float *C;
    long long ldc;
    const float *bias;  // per column, or null
    long long M;
    int N;
    long long K;
    int act;         // 0 none, 1 tanh
    int accumulate;  // atomic split-K accumulation into C (no bias / act)
    int kchunk;      // K per split (accumulate only; 0 = sized for ~1024 blocks)
};
// Bricks tile [x0, x0 + 32 nbx) x [y0, ..) x [z0, ..); cells at or beyond
// (x1, y1, z1) are outside the compressed region (zero / not written).
struct AeRegion {
    int x0, y0, z0, x1, y1, z1;
    int nbx, nby, nbz;
};
hipError_t launch_ae_gemm(const AeGemm &p, hipStream_t s);
hipError_t launch_ae_residual(const float *Y, const void *X, float *dY, long long n, float *loss,
                              hipStream_t s);
hipError_t launch_tanh_back(float *dH, const float *H, long long n, hipStream_t s);
hipError_t launch_colsum(const float *D, long long rows, int cols, float *db, hipStream_t s);
hipError_t launch_adam(float *w, const float *g, float *m, float *v, int n, float lr_t, float b1,
                       float b2, float eps, hipStream_t s);
hipError_t launch_gather_bricks(const void *src, const int *idx, void *dst, int n, hipStream_t s);
hipError_t launch_f32_to_f16(const float *a, void *b, long long n, hipStream_t s);
hipError_t launch_f16_to_f32(const void *a, float *b, long long n, hipStream_t s);
hipError_t launch_bricks(const float *F, const Grid &g, const AeRegion &r, void *out,
                         float *scales, hipStream_t s);
hipError_t launch_unbricks(const float *Y, const float *scales, const Grid &g, const AeRegion &r,
                           float *F, int b0, int nb, hipStream_t s);

}  // namespace rtmhip
