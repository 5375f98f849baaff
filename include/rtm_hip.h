/*
 * rtm_hip.h -- C ABI of the MI355X (gfx950) RTM propagation-and-imaging path.
 *
 * This is the drop-in boundary below the reference's operator interface
 * (class RTMPlatform, rtm-3d/include/RTMPlatform.hpp:22-114) and its CUDA
 * launchers (rtm-3d/kernel/gpu/rtmgpu.hpp:8-65).  Plain C types only: device
 * pointers are borrowed from the caller (the caller owns every buffer, as the
 * reference's grid objects own their device buffers), sizes are ints, every
 * call returns a status code (0 = success) and rtmhip_last_error() returns the
 * message of the last failure.
 *
 * Data layout is the reference's: [x][y][z], z fastest, fp32, extended grids
 * (nxe = nx + 2*blen, ...), offset = (x*nye + y)*nze + z
 * (rtm-3d/include/RTMGrid.hpp:217-221).
 *
 * All operators enqueue asynchronously on the library stream; copies to the
 * host and rtmhip_synchronize() wait for it.  Arithmetic follows the
 * reference CPU platform (rtm-3d/src/platform/host/RTMCPUPlatform.cpp)
 * operation for operation, so fp32 results are bit-identical to it.
 */
#ifndef RTM_HIP_H
#define RTM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define RTMHIP_OK               0
#define RTMHIP_ERR_INVALID      1  /* bad argument: sizes, pointers, order */
#define RTMHIP_ERR_HIP          2  /* HIP runtime error */
#define RTMHIP_ERR_DISPERSION   3  /* a NaN was produced by the stencil: the reference's
                                      "Numerical Dispersion" exception
                                      (RTMCPUPlatform.cpp:374-391) */
#define RTMHIP_ERR_NOT_INIT     4  /* rtmhip_init() not called */
#define RTMHIP_ERR_UNSUPPORTED  5  /* e.g. stencil order above 8 */

/* Receiver geometry in interior coordinates: the JSON keys
 * receiver_{start_x,start_y,depth_z,distance_x,distance_y,count_x,count_y}
 * (rtm-3d/src/seismic/RTMSeismic.cpp:4-30). */
typedef struct {
    int start_x, start_y, depth_z;
    int dist_x, dist_y;
    int count_x, count_y;
} rtmhip_rcv_geom;

/* One shot on one extended grid: everything the per-step operators of the
 * time loops need besides buffers.  Source coordinates are extended ones
 * (interior + blen), as RTMCPUPlatform::rtmApplySource computes them
 * (RTMCPUPlatform.cpp:103-113). */
typedef struct {
    int nxe, nye, nze;   /* extended grid */
    int blen;            /* absorbing border width */
    int order;           /* stencil order: 2, 4, 6 or 8 */
    int nt;              /* total time steps */
    int ntstep;          /* receiver-chunk length (divides nt) */
    int sxe, sye, sze;   /* source cell, extended coordinates */
    rtmhip_rcv_geom rcv; /* receiver geometry */
    int wave_format;     /* RTMHIP_WAVE_*: storage of the wavefields and UPB */
} rtmhip_shot;

/* Wavefield storage formats of the time loops (no reference counterpart:
 * the reference stores FP32 only).  Arithmetic is FP32 in every format; with
 * F16 or BF16 every wavefield pointer of the loop entry points (P, PP, Ps,
 * PPs, Pr, PPr and upb) refers to 2-byte elements and each wavefield write
 * is rounded to the format (nearest-even), exactly the reference's operator
 * sequence run on 16-bit grids.  Velocity, taper, Ricker, traces and the
 * image stay FP32.  16-bit formats support stencil order 8 only. */
#define RTMHIP_WAVE_F32  0
#define RTMHIP_WAVE_F16  1
#define RTMHIP_WAVE_BF16 2

/* ================= lifecycle =================
 * Replaces RTMGPUPlatform::initRTMPlatform / destroyRTMPlatform
 * (rtm-3d/src/platform/gpu/RTMGPUPlatform.cpp:17-57): selects the device,
 * creates the library stream.  Re-initialising with another device is an
 * error until rtmhip_finalize(). */
int rtmhip_init(int device);
int rtmhip_finalize(void);
const char *rtmhip_last_error(void);
int rtmhip_device_count(int *count);
int rtmhip_synchronize(void);
/* Library version as major*10000 + minor*100 + patch. */
int rtmhip_version(void);

/* ================= device-grid slot =================
 * Replaces GPUDeviceGrid::createDeviceBuffer / removeDeviceBuffer /
 * moveToDevice / moveFromDevice / devMemSet
 * (rtm-3d/src/grid/RTMGPUDeviceGrid.cpp:14-58,
 *  rtm-3d/include/RTMDeviceGrid.hpp:35-39). */
int rtmhip_alloc(void **dptr, size_t bytes);
int rtmhip_free(void *dptr);
int rtmhip_copy_to_device(void *dst, const void *src, size_t bytes);
int rtmhip_copy_from_device(void *dst, const void *src, size_t bytes);
int rtmhip_copy_device(void *dst, const void *src, size_t bytes);
int rtmhip_fill(float *dptr, size_t count, float value);
/* count 2-byte elements set to the bit pattern `bits` (16-bit wavefields;
 * 0 is +0.0 in both F16 and BF16).  No reference counterpart. */
int rtmhip_fill16(void *dptr, size_t count, unsigned short bits);

/* ================= per-step operators =================
 * One entry per RTMPlatform virtual (rtm-3d/include/RTMPlatform.hpp:28-102).
 * Host pointers are marked (host); everything else is a device pointer. */

/* rtmStep (RTMPlatform.hpp:50; CPU RTMCPUPlatform.cpp:325-395; CUDA
 * RTMSTEP_WRAPPER rtmgpu.hpp:8-11): PP <- 2P - PP + V2DT2*lap(P) on
 * [order/2, n-order/2)^3, in place.  coefs (host) is the [3][order+1] stencil
 * plane (RTMGrid.hpp:433-446).  A NaN makes a later rtmhip_synchronize()
 * return RTMHIP_ERR_DISPERSION. */
int rtmhip_step(const float *P, float *PP, const float *V2DT2, const float *coefs, int nxe,
                int nye, int nze, int order);

/* rtmTaperAllBorders / rtmTaperUpperBorders (RTMPlatform.hpp:36-43; CPU
 * taper3D RTMCPUPlatform.cpp:710-920, taperUpper3D :566-657; CUDA
 * RTMTAPER_WRAPPER rtmgpu.hpp:28-31).  taper: device table g[blen]
 * (RTMSeismic.hpp:294-305). */
int rtmhip_taper(float *P, const float *taper, int nxe, int nye, int nze, int blen,
                 int upper_only);

/* rtmApplySource (RTMPlatform.hpp:67; CPU RTMCPUPlatform.cpp:103-113; CUDA
 * RTMSOURCE_WRAPPER rtmgpu.hpp:22-26): PP[sxe][sye][sze] += value. */
int rtmhip_apply_source(float *PP, int nxe, int nye, int nze, int sxe, int sye, int sze,
                        float value);

/* rtmSaveReceiverData / rtmRestoreReceiverData (RTMPlatform.hpp:72-77; CPU
 * RTMCPUPlatform.cpp:114-151; CUDA RTMSEISM_WRAPPER rtmgpu.hpp:13-20).
 * rcv: one chunk [count_x][count_y][ntstep]; restore != 0 SETS the wavefield
 * (CPU semantics), restore == 0 samples it at time index lt. */
int rtmhip_receivers(float *PP, float *rcv, int nxe, int nye, int nze, int blen,
                     const rtmhip_rcv_geom *geom, int ntstep, int lt, int restore);

/* rtmSaveUpperBorder / rtmRestoreUpperBorder (RTMPlatform.hpp:82-88; CPU
 * RTMCPUPlatform.cpp:153-189; CUDA RTMWUPB_WRAPPER rtmgpu.hpp:33-34).
 * upb: [ntstep][nxe][nye][order/2] with 64-bit offsets; the layer is
 * z in [blen - order/2, blen). */
int rtmhip_upper_border(float *PP, float *upb, int nxe, int nye, int nze, int blen, int order,
                        int lt, int restore);

/* rtmImageCondition (RTMPlatform.hpp:62; CPU RTMCPUPlatform.cpp:513-540;
 * CUDA RTMIMG_WRAPPER rtmgpu.hpp:36-38): img += PS*PR on all x, y and
 * z in [blen, nze-blen) (CPU range). */
int rtmhip_image_condition(float *img, const float *PS, const float *PR, int nxe, int nye,
                           int nze, int blen);

/* rtmUpdateFreqContributions followed by rtmFreqDomainImageCondition for
 * every frequency of a group (RTMPlatform.hpp:93-102; CPU
 * RTMCPUPlatform.cpp:20-96; used by RTMFreqDomainKernel,
 * RTMFiniteDifferencesKernel_FreqDomain.cpp:301-315): for iw = iw0 + lw,
 * lw < nwstep, iw < nw, at time index st, on the interior cells
 *   sRe += kre[iw][st]*PS; sIm += kim[iw][st]*PS; rRe += kre[iw][st]*PR;
 *   rIm += kim[iw][st]*PR; img += w2[iw]*(sRe*rRe - sIm*rIm).
 * sre/sim/rre/rim: [nwstep][nx][ny][nz] (interior sizes); kre, kim:
 * [nw][nt]; w2: [nw] (squared angular steps). */
int rtmhip_freq_image(float *img, const float *PS, const float *PR, float *sre, float *sim,
                      float *rre, float *rim, const float *kre, const float *kim,
                      const float *w2, int nxe, int nye, int nze, int blen, int nwstep, int iw0,
                      int nw, int nt, int st);

/* The frequency-domain migration loop of one frequency group
 * (RTMFreqDomainKernel::rtmFreqDomainPropagation,
 * RTMFiniteDifferencesKernel_FreqDomain.cpp:220-357): for rt = it_hi down to
 * it_hi - nsteps + 1, with st = nt-1-rt and lt = rt % ntstep, both fields
 * stepped (source Ps/PPs, receiver Pr/PPr), taper3D of all four grids, the
 * Ricker sample st (device ricker[nt]) added to PPs, the receivers of
 * rcv_chunk at lt set into PPr, the contributions and image of the group as
 * rtmhip_freq_image at st, then the swaps (the pointers are swapped in
 * place).  FP32 wavefields.  Results equal that sequence of operator calls
 * bit for bit; the tapers are applied to the new fields' output when the
 * source and receivers lie off the borders (the default every-border taper
 * form, rtmhip_set_taper_fusion).  Calls may split the steps; the steps of
 * one call lie in one receiver chunk (else RTMHIP_ERR_INVALID). */
int rtmhip_freq_steps(const rtmhip_shot *shot, const float *V2DT2, const float *coefs,
                      const float *taper, const float *ricker, float **Ps, float **PPs,
                      float **Pr, float **PPr, const float *rcv_chunk, float *img, float *sre,
                      float *sim, float *rre, float *rim, const float *kre, const float *kim,
                      const float *w2, int nwstep, int iw0, int nw, int it_hi, int nsteps);

/* Image post-processing of the controller (RTMAcousticController.cpp:244-249,
 * RTMGrid.cpp:129-166 and :226-300): order-2 Laplacian filter with the
 * [3][3] plane fcoefs (host) followed by border removal; out is the
 * [nxe-2blen][nye-2blen][nze-2blen] image. */
int rtmhip_image_finish(const float *img, float *out, int nxe, int nye, int nze, int blen,
                        const float *fcoefs);

/* dst[i] <- dst[i] + src[i] (RTMCube::stackRegion, RTMGrid.cpp:195-223). */
int rtmhip_stack(float *dst, const float *src, size_t count);

/* ================= device-resident time loops =================
 * The reference's per-step loops of RTMFiniteDifferencesKernel /
 * RTMHBCKernel / RTMRBCKernel with no host synchronisation between steps.
 * The caller keeps two wavefield pointers per field and passes them by
 * address; the loops swap them exactly as RTMGRID_SWAP does (RTM.hpp:56-60),
 * so a run may be split into chunks of ntstep steps.  coefs: host plane as
 * for rtmhip_step; taper, ricker: device tables. */

/* Modeling, steps it0 .. it0+nsteps-1 of rtmAcousticFiniteDiffModeling
 * (rtm-3d/src/kernel/RTMFiniteDifferencesKernel.cpp:169-214): swap,
 * taperAll x2, step, source, save receivers into rcv_chunk at it % ntstep. */
int rtmhip_model_steps(const rtmhip_shot *shot, const float *V2DT2, const float *coefs,
                       const float *taper, const float *ricker, float **P, float **PP,
                       float *rcv_chunk, int it0, int nsteps);

/* Forward propagation of migration, steps it0 .. it0+nsteps-1:
 * hbc != 0: rtmHBCForward (RTMFiniteDifferencesKernel_HBC.cpp:140-185):
 *   swap, taperUpper x2, step, source, save the UPB layer of step it.
 *   hbc == 1: upb holds every step's layer (nt layers of nxe*nye*order/2
 *   elements; step it in layer it).  hbc == 2: upb holds one receiver
 *   chunk's layers (ntstep layers; step it in layer it % ntstep), the
 *   reference's per-chunk upper-border grid (RTMFiniteDifferencesKernel_HBC.cpp
 *   :34, :168; spilled to a file after each chunk, :188-195, and loaded back
 *   before the backward pass reaches that chunk, :261-277); the steps of one
 *   call must then lie in one receiver chunk.
 * hbc == 0: rtmRBCForward (RTMFiniteDifferencesKernel_RBC.cpp:96-126). */
int rtmhip_forward_steps(const rtmhip_shot *shot, int hbc, const float *V2DT2,
                         const float *coefs, const float *taper, const float *ricker, float **P,
                         float **PP, float *upb, int it0, int nsteps);

/* Backward propagation with imaging, it from it_hi down to it_hi-nsteps+1
 * (rtmHBCBackward RTMFiniteDifferencesKernel_HBC.cpp:261-354,
 * rtmRBCBackward RTMFiniteDifferencesKernel_RBC.cpp:201-266).  On the first
 * call (it_hi = nt-1) *Ps and *PPs must be the forward's final *P and *PP
 * (the reference's snap0/snap1); *Pr, *PPr and img must be zero.  rcv_chunk
 * holds the chunk containing it_hi.  With fused != 0 the receiver step,
 * receiver injection and imaging condition run as one pass; with fused == 2
 * (FP32 wavefields, order 8) the source field's backward step joins that
 * pass for every it < nt-2, so V is read once and the image takes the new
 * source value from registers.  Results are identical in all modes.  hbc
 * selects the UPB layout as for rtmhip_forward_steps (2: upb holds the
 * layers of the receiver chunk being processed). */
int rtmhip_backward_steps(const rtmhip_shot *shot, int hbc, const float *V2DT2,
                          const float *coefs, const float *taper, float **Ps, float **PPs,
                          float **Pr, float **PPr, const float *upb, const float *rcv_chunk,
                          float *img, int it_hi, int nsteps, int fused);

/* ================= snapshot RTM =================
 * No reference counterpart as a loop: the reference recovers the source
 * field for imaging by HBC reconstruction (rtmHBCBackward,
 * RTMFiniteDifferencesKernel_HBC.cpp:261-354).  Here the forward pass stores
 * the source wavefield of every k-th step -- the steps it = nt-1, nt-1-k,
 * nt-1-2k, ... ("snapshot steps"; snapshot j is step nt-1-j*k) -- over the
 * imaging range (all x, y; z in [blen, nze-blen)) as [j][nxe][nye][nze-2*blen]
 * elements of snap_format, and the backward pass propagates only the
 * receiver field and applies the imaging condition at the snapshot steps:
 * img += k * (snapshot * PR).  With k = 1 and F32 snapshots this is the
 * reference's imaging condition (RTMCPUPlatform.cpp:513-540) on the forward
 * source field ("exact-snapshot RTM"); k > 1 images every k-th step, weighted
 * so that the image approximates the every-step one.  The forward step
 * sequence is rtmhip_forward_steps' with hbc = 1 minus the upper-border
 * history; the receiver half of the backward pass is rtmhip_backward_steps'
 * (hbc = 1).  Calls may split the time range into chunks like those loops.
 * wave_format may be F32, F16 or BF16 (the snapshot format is independent). */
#define RTMHIP_SNAP_F32 0  /* FP32 snapshots */
#define RTMHIP_SNAP_F16 1  /* FP16 snapshots (nearest-even), FP32 imaging arithmetic */

/* Size in bytes of the snapshot buffer of one shot. */
int rtmhip_snap_buffer_bytes(const rtmhip_shot *shot, int snap_format, int k, size_t *bytes);
int rtmhip_snap_forward_steps(const rtmhip_shot *shot, const float *V2DT2, const float *coefs,
                              const float *taper, const float *ricker, float **P, float **PP,
                              void *snaps, int snap_format, int k, int it0, int nsteps);
/* it from it_hi down to it_hi-nsteps+1; *Pr, *PPr and img zero on the first
 * call (it_hi = nt-1); rcv_chunk holds the chunk containing it_hi. */
int rtmhip_snap_backward_steps(const rtmhip_shot *shot, const float *V2DT2, const float *coefs,
                               const float *taper, float **Pr, float **PPr, const void *snaps,
                               int snap_format, int k, const float *rcv_chunk, float *img,
                               int it_hi, int nsteps);

/* Snapshot RTM with transform-coded snapshots: the loops of
 * rtmhip_snap_forward_steps / rtmhip_snap_backward_steps (same snapshot
 * steps and imaging) with each snapshot of the imaging range kept as the
 * largest coefficients of its 3D DCT-II over 8 x 8 x 32 bricks: at most
 * K = nxe*nye*(nze-2*blen) / keep_div coefficients per snapshot (one
 * power-of-two threshold for the whole snapshot, so energetic bricks keep
 * more), each stored as 32 bits (index in the brick, FP16 value over the
 * brick's largest magnitude), plus a 12-byte table entry per brick.  With
 * keep_div = 32 the store is about 16x smaller than FP16 snapshots.  The
 * backward pass decodes each brick in on-chip memory and applies
 * img += k * (v * PR) directly.  Any wave format.  Results are FP32
 * arithmetic (fused multiply-adds in the transforms), not bit-identical to
 * a reference operator: the image error against exact snapshots is
 * measured.  rtmhip_snap_dct_sizes gives the byte sizes of the store (all
 * snapshots, after an 8 KB header holding the basis) and of the work buffer
 * (any output may be NULL).
 * basis (host, forward and roundtrip calls): NULL for the DCT-II, or a
 * learned separable basis -- 32x32 along z, then 8x8 along x, then 8x8 along
 * y, row-major, each row a basis vector, rows orthonormal to 1e-4 (else
 * RTMHIP_ERR_INVALID), the same on every forward call of one shot; the
 * store keeps it for the backward pass. */
int rtmhip_snap_dct_sizes(const rtmhip_shot *shot, int k, int keep_div, size_t *store_bytes,
                          size_t *work_bytes);
int rtmhip_snap_dct_forward_steps(const rtmhip_shot *shot, const float *V2DT2,
                                  const float *coefs, const float *taper, const float *ricker,
                                  float **P, float **PP, void *store, void *work,
                                  const float *basis, int k, int keep_div, int it0, int nsteps);
int rtmhip_snap_dct_backward_steps(const rtmhip_shot *shot, const float *V2DT2,
                                   const float *coefs, const float *taper, float **Pr,
                                   float **PPr, const void *store, int k, int keep_div,
                                   const float *rcv_chunk, float *img, int it_hi, int nsteps);
/* The coder alone: encodes the imaging range (z in [blen, nze-blen)) of one
 * field of the given wave format into store (one snapshot's slot: the size
 * rtmhip_snap_dct_sizes gives for nt = 1) and decodes it into the same
 * cells of out (FP32 grid); *kept (host, optional) receives the number of
 * coefficients kept. */
int rtmhip_dct_roundtrip(const void *field, int wave_format, int nxe, int nye, int nze,
                         int blen, int keep_div, void *store, void *work, const float *basis,
                         float *out, unsigned *kept);

/* ================= batches of shots =================
 * No reference counterpart: the reference migrates one shot at a time per
 * process (rtm-3d/src/controller/RTMAcousticController.cpp:141-200).  A batch
 * advances many HBC shots, each on its own extended grid (uneven apertures),
 * in lockstep: every step is one stencil launch per kernel class covering
 * all shots plus one launch of the small per-step operators (frame taper,
 * source, UPB, receivers, imaging off the stencil's cells) of all shots, and
 * runs of steps replay a captured HIP graph.  Each shot's results are
 * bit-identical to rtmhip_forward_steps (hbc = 1) and rtmhip_backward_steps
 * (hbc = 1, fused = 1) on that shot alone.
 *
 * Every shot: stencil order 8, the same blen, nt, wave_format, coefficient
 * plane and taper table; ntstep == nt (rcv holds the whole record
 * [count_x][count_y][nt]); upb holds all nt layers.  The caller owns every
 * buffer; the batch borrows them until rtmhip_batch_destroy.  fs and fr are
 * the source and receiver wavefield pairs, (fs[0], fs[1]) taken as
 * (P, PP) at creation; rtmhip_batch_fields returns the current roles.  A
 * forward run (rtmhip_batch_forward_steps from it0 = 0 to nt-1) leaves the
 * source pair as the reference's snap0 / snap1 for the backward run
 * (rtmhip_batch_backward_steps from it_hi = nt-1 down), which needs fr and
 * img zeroed.  Calls may split the time range into chunks. */
typedef struct {
    rtmhip_shot shot;
    const float *V2DT2;    /* v^2 dt^2, extended grid */
    const float *ricker;   /* [nt] source samples */
    float *fs[2];          /* source wavefield pair */
    float *fr[2];          /* receiver wavefield pair */
    float *upb;            /* [nt][nxe][nye][order/2] */
    const float *rcv;      /* [count_x][count_y][nt] */
    float *img;            /* extended-grid image (accumulated) */
} rtmhip_batch_shot;

typedef struct rtmhip_batch rtmhip_batch;

/* coefs (host): the shared [3][9] stencil plane; taper: device table g[blen]. */
int rtmhip_batch_create(rtmhip_batch **batch, const rtmhip_batch_shot *shots, int nshots,
                        const float *coefs, const float *taper);
int rtmhip_batch_destroy(rtmhip_batch *batch);
int rtmhip_batch_forward_steps(rtmhip_batch *batch, int it0, int nsteps);
int rtmhip_batch_backward_steps(rtmhip_batch *batch, int it_hi, int nsteps);
/* Current (P, PP) of shot i's source pair and (Pr, PPr) of its receiver
 * pair; any output may be NULL. */
int rtmhip_batch_fields(const rtmhip_batch *batch, int i, float **P, float **PP, float **Pr,
                        float **PPr);
/* Stencil launches of one kernel class (0 backward source step, 1 receiver
 * step with imaging, 2 forward tapered step): mode 0 = one launch per shot
 * with the class tiling of rtmhip_set_tiling; 6 or 8 = one launch for every
 * shot (barrier-free 64-lane waves with 6 or 8 rows each).  Default: mode 0
 * for every class. */
int rtmhip_batch_set_stencil(rtmhip_batch *batch, int kernel_class, int mode);
/* Steps per captured graph: 0 = launch every step (no graph), else an even
 * count (default 8). */
int rtmhip_batch_set_graph(rtmhip_batch *batch, int steps);

/* No reference counterpart.  For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): nsteps batched HBC steps of the shots described by `shots`
 * alone, with no batch object, no graph and no device allocation -- forward
 * (direction = +1, steps it .. it+nsteps-1, as rtmhip_batch_forward_steps)
 * or backward (direction = -1, it down to it-nsteps+1, as
 * rtmhip_batch_backward_steps).  Results are bit-identical to a batch
 * created from the same shots and run over the same chunks.  On entry
 * (fs[0], fs[1]) and (fr[0], fr[1]) are each pair's current (P, PP); on
 * return they hold the new roles.  The shot table is staged in a fixed device
 * array of the library holding at most 64 shots (one per GPU, kept until the
 * process ends, holding the last call's table); more shots are rejected with
 * RTMHIP_ERR_INVALID.  Calls run in order on the library stream; calls from
 * several host threads are serialised. */
int rtmhip_batch_steps_noalloc(rtmhip_batch_shot *shots, int nshots, const float *coefs,
                               const float *taper, int direction, int it, int nsteps);

/* ================= learned snapshot compression =================
 * No reference counterpart (the reference keeps the source field by HBC
 * reconstruction or by exact snapshots).  A 4-layer patch-convolution
 * autoencoder on 32^3 bricks of a source-wavefield snapshot:
 *   L1 2^3 patches 1->8 channels, tanh;  L2 4^3 patches 8->32 (the code);
 *   L3, L4 the transposed layers back to 32^3, tanh after L3.
 * A brick is stored as FP16 values normalised by its largest magnitude
 * (scale, FP32), in the library's brick element order; its code is 2048
 * FP16 values: 16x smaller than the FP16 brick (32x than FP32).  Every
 * layer runs on the matrix cores (FP16 operands, FP32 accumulation), in
 * inference and in training (forward, backward, Adam), on the device.
 *
 * A region [x0,x1) x [y0,y1) x [z0,z1) of an extended grid is cut into
 * ceil(dx/32) x ceil(dy/32) x ceil(dz/32) bricks, x slowest, z fastest;
 * cells past the region's end are zero in a brick and never written back. */
typedef struct rtmhip_ae rtmhip_ae;

int rtmhip_ae_param_count(void);
/* Glorot-uniform weights from seed; max_bricks (1..4096) sizes the
 * workspace (bricks per training batch or per encode/decode pass). */
int rtmhip_ae_create(rtmhip_ae **ae, int max_bricks, unsigned long long seed);
int rtmhip_ae_destroy(rtmhip_ae *ae);
/* Parameters (which = 0) or the last gradient (which = 1) to the host; the
 * layout is W1[8][8] b1[8] W2[512][32] b2[32] W3[32][512] b3[512]
 * W4[8][8] b4[8], W[k][n] row-major. */
int rtmhip_ae_get_params(const rtmhip_ae *ae, float *host, int which);
int rtmhip_ae_set_params(rtmhip_ae *ae, const float *host);
/* Gradient of 0.5 * sum (decode(encode(x)) - x)^2 over n FP16 bricks (device)
 * into the gradient buffer; *loss (host) receives that half sum. */
int rtmhip_ae_gradient(rtmhip_ae *ae, const void *bricks, int n, float *loss);
/* iters Adam steps (lr; beta 0.9, 0.999; eps 1e-8) on batches of `batch`
 * bricks drawn without replacement from nbricks FP16 bricks (device);
 * losses (host, optional): the mean squared error of every batch. */
int rtmhip_ae_train(rtmhip_ae *ae, const void *bricks, int nbricks, int iters, int batch,
                    float lr, unsigned long long seed, float *losses);
/* Cut a region of an FP32 field into normalised FP16 bricks + scales. */
int rtmhip_ae_bricks(const float *field, int nxe, int nye, int nze, int x0, int y0, int z0,
                     int x1, int y1, int z1, void *bricks, float *scales);
/* n FP16 bricks -> n x 2048 FP16 codes. */
int rtmhip_ae_encode(rtmhip_ae *ae, const void *bricks, int n, void *codes);
/* Codes + scales of a region's n bricks -> the region's cells of field. */
int rtmhip_ae_decode(rtmhip_ae *ae, const void *codes, const float *scales, int n, float *field,
                     int nxe, int nye, int nze, int x0, int y0, int z0, int x1, int y1, int z1);
/* Snapshot RTM with coded snapshots: the loops of rtmhip_snap_forward_steps
 * and rtmhip_snap_backward_steps (same snapshot steps, same imaging
 * img += k * (snapshot * PR)) with each snapshot's imaging region (all x, y;
 * z in [blen, nze-blen)) kept as codes + scales instead of an F32/F16 copy:
 * the forward pass cuts it into bricks in `work` and encodes them into slot
 * j of `codes` ([nsnap][bricks][2048] FP16) and `scales`
 * ([nsnap][bricks] FP32); the backward pass decodes slot j into `work` as a
 * compact [nxe][nye][nze-2*blen] FP32 field and images with it.  The result
 * equals rtmhip_snap_backward_steps (F32 snapshots) run on the decoded
 * snapshots.  FP32 wavefields only (RTMHIP_ERR_UNSUPPORTED otherwise).
 * rtmhip_snap_ae_sizes gives the bricks per snapshot and the byte sizes of
 * the three buffers (any output may be NULL).  The autoencoder's workspace is
 * used by both passes: one shot at a time per autoencoder. */
int rtmhip_snap_ae_sizes(const rtmhip_shot *shot, int k, int *bricks, size_t *code_bytes,
                         size_t *scale_bytes, size_t *work_bytes);
int rtmhip_snap_ae_forward_steps(const rtmhip_shot *shot, const float *V2DT2, const float *coefs,
                                 const float *taper, const float *ricker, float **P, float **PP,
                                 rtmhip_ae *ae, void *codes, float *scales, void *work, int k,
                                 int it0, int nsteps);
int rtmhip_snap_ae_backward_steps(const rtmhip_shot *shot, const float *V2DT2,
                                  const float *coefs, const float *taper, float **Pr, float **PPr,
                                  rtmhip_ae *ae, const void *codes, const float *scales,
                                  void *work, int k, const float *rcv_chunk, float *img,
                                  int it_hi, int nsteps);
/* One matrix-core GEMM of the autoencoder, exposed for tests:
 * C(m,n) (+)= sum_k A(m,k) B(k,n) (+ bias[n], tanh if act), with
 * A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn], FP16 (x_half != 0)
 * or FP32 in memory; accumulate != 0 adds split-K partials atomically. */
int rtmhip_ae_gemm_test(const void *A, long long sam, long long sak, int a_half, const void *B,
                        long long sbk, long long sbn, int b_half, float *C, long long ldc,
                        const float *bias, long long M, int N, long long K, int act,
                        int accumulate);

/* ================= tuning and measurement =================
 * Stencil kernel classes: 0 = plain step (rtmhip_step, backward source
 * reconstruction, RBC), 1 = receiver step fused with receiver injection and
 * imaging, 2 = HBC forward step with the upper taper fused, 3 = source and
 * receiver backward steps in one pass (fused == 2; shape 64x1x4).  No reference
 * counterpart: the reference's launch shape is fixed (10x10x10 blocks,
 * RTMGPUPlatform.cpp:176-185).
 *
 * rtmhip_set_tiling selects the tile shape of one class: tz lanes along z,
 * by thread rows, ry rows per thread, xchunk planes marched per block (0 =
 * sized from the device's resident-block capacity).  Compiled shapes:
 *   tz = 64: tiles 64x4x4, 64x4x2, 64x8x2; barrier-free waves 64x1x6, 64x1x8;
 *   tz = 32: tiles 32x8x4, 32x8x2;
 *   tz = 1 (order 8): full-z-row blocks 4x4 (nze <= 256), 2x6 and 2x7
 *           (nze <= 384);
 *   tz = 2 (order 8, even nze <= 384): full rows, two z per lane, 4x3;
 *   tz = 128 (order 8): barrier-free waves, two z per lane, 128x1x4;
 *           128x1x6 and 128x1x8: one-z 64x1x6 / 64x1x8 waves on the z
 *           tiles the upper taper reaches and two-z waves below (odd nze or
 *           strides: one-z waves throughout).
 * tz = 0 restores the default: full-z-row blocks where they fit, else (FP32,
 * order 8) 64x1x8 waves, else 64x4x4; a full-row request that does not
 * cover nze takes the same fallback.  16-bit wavefields run the 64x1xry
 * and 128x1x4 waves and the full rows 4x4 / 2x6 on request and otherwise
 * their defaults (128x1x4 waves for untapered steps where nze and the
 * strides are even, else 64x1x6 waves for the imaging step and 64x1x8 waves
 * otherwise). */
int rtmhip_set_tiling(int kernel_class, int tz, int by, int ry, int xchunk);
int rtmhip_get_tiling(int kernel_class, int *tz, int *by, int *ry, int *xchunk);

/* How the tapers reach the loops, per wave format; results are identical in
 * every mode, only the speed differs.
 * HBC upper tapers (taperUpper3D of both fields before each step,
 * RTMFiniteDifferencesKernel_HBC.cpp:160-165 and :311-316) in the single-shot
 * loops (rtmhip_forward_steps, rtmhip_backward_steps with fused = 1,
 * rtmhip_snap_*): mode 1 fuses them into the stencil (tapered on load);
 * mode 0 runs them as one in-place pass over both fields' upper slabs
 * followed by the untapered stencil; mode 4 applies the forward loops' tapers
 * to each step's output (the step stores the tapered value its successor
 * reads as P and tapers only the centre PP; used when the source lies at
 * z >= blen) and fuses the backward loops' on load.
 * Every-border tapers of rtmhip_model_steps (taper3D of both fields before
 * each step, RTMFiniteDifferencesKernel.cpp:180-181; stencil order 8): on
 * each step's output with modes 3 and 4 (used when the source and receivers
 * lie off the borders, else the separate pass runs), fused on load with
 * mode 2 (upper tapers as mode 1), one in-place pass with modes 0 and 1.
 * Mode -1 restores the default: mode 4 for F32; for F16 and BF16 the upper
 * tapers in separate passes and the every-border tapers on the output.
 * Batches and fused = 2 always fuse the upper tapers on load.
 * rtmhip_get_taper_fusion returns the last mode set (-1 for the default). */
int rtmhip_set_taper_fusion(int wave_format, int mode);
int rtmhip_get_taper_fusion(int wave_format, int *fused);

/* When enabled, every launch of the stencil kernels is bracketed by HIP
 * events on the library stream; rtmhip_timing_read() synchronises and returns
 * the number of launches and their summed duration for one kernel class,
 * then clears that class's record. */
int rtmhip_timing_enable(int on);
int rtmhip_timing_read(int kernel_class, long long *launches, double *total_ms);

#ifdef __cplusplus
}
#endif
#endif
