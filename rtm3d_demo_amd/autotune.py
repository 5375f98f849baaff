"""Tile-shape autotuner for the stencil kernels.

For a given grid, every compiled tile shape is timed for the three stencil
kernel classes (0 plain step, 1 fused receiver step, 2 tapered HBC forward
step) with HIP events around each launch, by running a few steps of the
forward and fused backward loops on the caller's buffers; the fastest shape
per class is installed with rtmhip_set_tiling.  Results are identical for
every shape (tested), so tuning only changes speed.
"""
import numpy as np


def tune(hip, shot, V, coefs, g, src, F, R, upb, rcv, img, steps=6, rounds=2, log=None):
    """Returns {class: ((tz, by, ry), avg_ms)} and installs the winners.

    The buffers are scratch for the duration of the call (their contents are
    overwritten); `shot` must allow `steps` time steps."""
    best = {}
    nt = shot.nt
    for tiling in hip.TILINGS:
        for cls in range(3):
            hip.set_tiling(cls, *tiling, 0)
        # one untimed pass to settle clocks and caches
        hip.forward_steps(shot, True, V, coefs, g, src, F, upb, 0, steps)
        hip.synchronize()
        hip.timing_enable(True)
        for _ in range(rounds):
            hip.forward_steps(shot, True, V, coefs, g, src, F, upb, 0, steps)
            hip.backward_steps(shot, True, V, coefs, g, F, R, upb, rcv, img, nt - 1, steps, True)
        res = [hip.timing_read(c) for c in range(3)]
        hip.timing_enable(False)
        for cls, (n, ms) in enumerate(res):
            if n == 0:
                continue
            avg = ms / n
            if log:
                log("tile %s class %d: %.4f ms" % (tiling, cls, avg))
            if cls not in best or avg < best[cls][1]:
                best[cls] = (tiling, avg)
    # With the forward upper taper on the output (the FP32 default) every
    # class-2 shape but the one-z 64x1x6 / 64x1x8 waves launches the same
    # two-z 128x1x4 kernel: record that shape
    fmt = {0: "f32", 1: "f16", 2: "bf16"}.get(getattr(shot, "wave_format", 0), "f32")
    if 2 in best and hip.get_taper_fusion(fmt) in (-1, 4) and fmt == "f32":
        t2 = tuple(best[2][0])
        if not (t2[0] == 64 and t2[1] == 1 and t2[2] in (6, 8)):
            best[2] = ((128, 1, 4), best[2][1])
    for cls in range(3):
        if cls in best:
            hip.set_tiling(cls, *best[cls][0], 0)
        else:
            hip.set_tiling(cls, 0)
    for f in (F.a, F.b, R.a, R.b, img):
        f.zero()
    return best


BATCH_SHOT_TILINGS = [(1, 4, 4), (1, 2, 6), (1, 2, 7), (2, 4, 3), (64, 1, 8), (128, 1, 4)]


def tune_batch(hip, batch, nt, steps=17, log=None):
    """Autotunes a batch's stencil launches per kernel class: one launch per
    shot with each candidate tile shape (a full-row shape that does not cover
    a shot's nz runs that shot with the fallback), or one launch for every
    shot (barrier-free waves, 6 or 8 rows).  Each candidate is timed over
    `steps` batched forward (class 2) or backward (classes 0, 1) steps: the
    first step of each call runs eagerly, the rest replay the batch's captured
    graphs (steps = 17 is two replays of the default 8-step graph), so the
    timing covers the graph form the batch runs.

    Installs the fastest form per class on the batch, and the fastest
    per-shot tile shape per class with set_tiling (so single-shot runs of the
    same grids use it too).  Returns {class: {"mode", "tile", "ms_per_step",
    "tile_ms_per_step"}}.  The batch's buffers are scratch for the duration
    of the call."""
    import time

    out = {}
    for cls, fwd in ((2, True), (0, False), (1, False)):
        def run():
            if fwd:
                batch.forward_steps(0, steps)
            else:
                batch.backward_steps(nt - 3, steps)

        def timed():
            run()  # untimed: captures this candidate's graphs
            hip.synchronize()
            t = time.perf_counter()
            for _ in range(2):
                run()
            hip.synchronize()
            return (time.perf_counter() - t) / (2 * steps) * 1e3

        best_tile, best_mode = None, None
        for tiling in BATCH_SHOT_TILINGS:
            batch.set_stencil(cls, 0)
            hip.set_tiling(cls, *tiling, 0)
            ms = timed()
            if log:
                log("batch class %d per-shot %s: %.3f ms/step" % (cls, tiling, ms))
            if best_tile is None or ms < best_tile[1]:
                best_tile = (tiling, ms)
        hip.set_tiling(cls, *best_tile[0], 0)
        best_mode = (0, best_tile[1])
        for mode in (6, 8):
            batch.set_stencil(cls, mode)
            ms = timed()
            if log:
                log("batch class %d one launch, %d rows: %.3f ms/step" % (cls, mode, ms))
            if ms < best_mode[1]:
                best_mode = (mode, ms)
        batch.set_stencil(cls, best_mode[0])
        out[cls] = {"mode": best_mode[0], "ms_per_step": round(best_mode[1], 3),
                    "tile": "%dx%dx%d" % best_tile[0],
                    "tile_ms_per_step": round(best_tile[1], 3)}
    return out


def describe(best):
    return {str(c): {"tile": "%dx%dx%d" % t, "avg_ms": round(ms, 4)} for c, (t, ms) in
            sorted(best.items())}


__all__ = ["tune", "tune_batch", "describe", "np"]
