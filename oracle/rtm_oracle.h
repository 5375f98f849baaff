/*
 * rtm_oracle.h -- CPU oracle for the RTM propagation-and-imaging path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: the parity tests,
 * the smoke check and the CPU-baseline leg of the benchmark may load it; the
 * product path (librtm3d_hip.so and the rtm3d driver) never links, loads or
 * calls it, and there is no CPU fallback anywhere in the product.
 *
 * Two families of functions:
 *   - orc_*_f32: an FP32 restatement of the reference CPU platform with the
 *     same loop order, the same operation order, no FMA and no flush-to-zero,
 *     so that results are bit-identical to the reference program.
 *   - orc_*_f64: the same algorithm carried out in double precision from the
 *     same FP32 model, used as the accuracy yardstick for every dtype variant.
 *   - orc_*_f16, orc_*_bf16: the FP32 algorithm with 16-bit wavefield
 *     storage (each wavefield write rounded to the format), the bit-exact
 *     yardstick of the 16-bit GPU path.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the reference tree; rtm-3d/ = source/rtm-3d/).
 *
 * Pinning (what checks each function independently of its own formula):
 *   orc_stencil_coefs_f32   exact rationals; sum of taps = 0; 2nd moment = 2;
 *                           fixtures of the reference program.
 *   orc_ricker_f32          closed form in double, fixtures.
 *   orc_taper_table_f32     closed form Gaussian.
 *   orc_glibc_rand_*        libc rand() itself (same glibc), first 10^5 draws.
 *   orc_borders_*_f32       reference-produced extended velocity files (vpe).
 *   orc_step_*              quadratic/cubic exactness, impulse symmetry,
 *                           brute-force double evaluation on tiny grids,
 *                           reference seismograms.
 *   orc_taper_*             brute-force index enumeration incl. the
 *                           front-right quirk; reference seismograms.
 *   orc_receivers/upb/image save/restore round trips; reference outputs.
 *   orc_*_shot / migrate    reference seismogram and image fixtures
 *                           (bit-exact), HBC reversibility.
 *   orc_f32_to_f16 etc.     numpy float16 casts (all 65536 half values one
 *                           way, 10^7 random and edge floats the other);
 *                           torch bfloat16 casts.
 *   orc_*_f16 / _bf16       equal to the f32 build when every wavefield value
 *                           is exactly representable (integers, dyadic
 *                           taps); error vs f64 within the format's bound.
 */
#ifndef RTM_ORACLE_H
#define RTM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Receiver geometry in interior (non-extended) coordinates, as the JSON
 * keys receiver_{start_x,start_y,depth_z,distance_x,distance_y,count_x,count_y}
 * (rtm-3d/src/seismic/RTMSeismic.cpp:4-30). */
typedef struct {
    int start_x, start_y, depth_z;
    int dist_x, dist_y;
    int count_x, count_y;
} orc_rcv_geom;

/* ---------------- host-side set-up (float, reference semantics) -------- */

/* Laplacian taps for an order (2,4,6,8 closed forms; other even orders use
 * the windowed series); rtm-3d/src/misc/Misc.cpp:106-176. out[order+1]. */
void orc_laplacian_taps_f32(int order, float *out);
/* Stencil plane [3][order+1] = (1/d)^2 * taps, rtm-3d/include/RTMGrid.hpp:433-446. */
void orc_stencil_coefs_f32(int order, float dx, float dy, float dz, float *out);
/* Ricker samples s[it], rtm-3d/include/RTMSeismic.hpp:135-144,191-198. */
void orc_ricker_f32(int nt, float dt, float fpeak, float *out);
/* Gaussian taper g[i], i<blen, rtm-3d/include/RTMSeismic.hpp:294-305. */
void orc_taper_table_f32(int blen, float factor, float *out);
/* Frequency-domain migration tables (w^2, cos / sin kernels [nw][nt]),
 * RTMFiniteDifferencesKernel_FreqDomain.cpp:17-52, 125-152; returns nw. */
int orc_freq_tables(int nt, float dt, float fmin, float fmax, float *w2, float *kre, float *kim,
                    int cap);

/* glibc TYPE_3 additive-feedback generator (the algorithm behind rand()). */
typedef struct { int32_t r[34]; int32_t tbl[31]; int f, b; } orc_rand_state;
void orc_glibc_rand_seed(orc_rand_state *st, unsigned seed);
int  orc_glibc_rand_next(orc_rand_state *st);

/* RBC velocity randomiser, rtm-3d/src/misc/Misc.cpp:91-102. */
float orc_randlinearval(float v, int k, int nkb, orc_rand_state *st);

/* Interior copy of a [nx][ny][nz] model into an extended grid
 * (rtm-3d/src/grid/RTMGrid.cpp:47-86); border cells are zeroed. */
void orc_extend_borders_f32(const float *v, int nx, int ny, int nz, int blen, float *ve);
/* Border fills on an extended [nxe][nye][nze] grid, RTMGrid.cpp:371-696. */
void orc_borders_abc_f32(float *ve, int nxe, int nye, int nze, int blen);
void orc_borders_rbc_f32(float *ve, int nxe, int nye, int nze, int blen, orc_rand_state *st);
void orc_borders_hbc_f32(float *ve, int nxe, int nye, int nze, int blen, orc_rand_state *st);
/* v <- (v*v)*(dt*dt), rtm-3d/src/controller/RTMController.cpp:110-111. */
void orc_v2dt2_f32(float *v, size_t n, float dt);

/* Double-precision set-up for the f64 oracle (ideal values, not a restatement). */
void orc_stencil_coefs_f64(int order, double dx, double dy, double dz, double *out);
void orc_ricker_f64(int nt, double dt, double fpeak, double *out);
void orc_taper_table_f64(int blen, double factor, double *out);
void orc_v2dt2_f64(const float *v, size_t n, double dt, double *out);

/* ---------------- 16-bit storage formats -------------------------------
 * binary16 and bfloat16 <-> binary32, nearest-even, subnormals kept. */
uint16_t orc_f32_to_f16(float f);
float    orc_f16_to_f32(uint16_t h);
uint16_t orc_f32_to_bf16(float f);
float    orc_bf16_to_f32(uint16_t b);
void orc_to_f16(const float *in, uint16_t *out, size_t n);
void orc_from_f16(const uint16_t *in, float *out, size_t n);
void orc_to_bf16(const float *in, uint16_t *out, size_t n);
void orc_from_bf16(const uint16_t *in, float *out, size_t n);

/* ---------------- per-step operators -----------------------------------
 * R = arithmetic type, W = wavefield storage type:
 *   f32 (float, float), f64 (double, double): the reference algorithm;
 *   f16, bf16 (float, uint16_t): float arithmetic, 16-bit wavefields, every
 *   wavefield write rounded to the format (nearest-even). */
#define ORC_DECLARE_OPS(R, W, S)                                                       \
int  orc_step_##S(const W *P, W *PP, const R *V2, const R *coefs,                      \
                  int nxe, int nye, int nze, int order);                               \
void orc_taper_all_##S(W *P, const R *g, int nxe, int nye, int nze, int blen);         \
void orc_taper_upper_##S(W *P, const R *g, int nxe, int nye, int nze, int blen);       \
void orc_apply_source_##S(W *PP, int nxe, int nye, int nze, int sxe, int sye, int sze, \
                          R val);                                                      \
void orc_receivers_##S(W *PP, R *rcv, int nxe, int nye, int nze, int blen,             \
                       const orc_rcv_geom *g, int ntstep, int lt, int restore);        \
void orc_upb_##S(W *PP, W *upb, int nxe, int nye, int nze, int blen, int order,        \
                 int lt, int restore);                                                 \
void orc_image_##S(R *img, const W *PS, const W *PR, int nxe, int nye, int nze,        \
                   int blen);                                                          \
void orc_image_finish_##S(const R *img, R *out, int nxe, int nye, int nze, int blen,   \
                          const R *fcoefs);                                            \
int  orc_model_shot_##S(const R *v2, const R *coefs, const R *g, const R *src,         \
                        int nxe, int nye, int nze, int blen, int order, int nt,        \
                        int ntstep, int sx, int sy, int sz, const orc_rcv_geom *rg,    \
                        R *rcv);                                                       \
int  orc_migrate_shot_##S(const R *v2, const R *coefs, const R *g, const R *src,       \
                          int nxe, int nye, int nze, int blen, int order, int nt,      \
                          int ntstep, int sx, int sy, int sz, const orc_rcv_geom *rg,  \
                          const R *rcv, int hbc, R *img, W *fwd_last);                  \
void orc_freq_image_##S(R *img, const W *PS, const W *PR, R *sre, R *sim, R *rre,       \
                        R *rim, const R *kre, const R *kim, const R *w2, int nxe,      \
                        int nye, int nze, int blen, int nwstep, int iw0, int nw,       \
                        int nt, int st);                                               \
int  orc_fmig_shot_##S(const R *v2, const R *coefs, const R *g, const R *src,          \
                       int nxe, int nye, int nze, int blen, int order, int nt,         \
                       int ntstep, int sx, int sy, int sz, const orc_rcv_geom *rg,     \
                       const R *rcv, const R *kre, const R *kim, const R *w2, int nw,  \
                       int nwstep, R *img);

ORC_DECLARE_OPS(float, float, f32)
ORC_DECLARE_OPS(double, double, f64)
ORC_DECLARE_OPS(float, uint16_t, f16)
ORC_DECLARE_OPS(float, uint16_t, bf16)

#ifdef __cplusplus
}
#endif
#endif
