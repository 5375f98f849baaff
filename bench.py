#!/usr/bin/env python3
"""Benchmark of the RTM propagation-and-imaging path on one MI355X.

Headline (one "step" = one complete FP32 HBC RTM shot: forward propagation
with upper-border history, backward source reconstruction, receiver
back-propagation and cross-correlation imaging):
  512x512x256 salt model at 12.5 m (v(z) gradient, seeded smoothed +-5 %
  perturbation, seeded ellipsoidal salt at 4500 m/s), HBC border 40
  (extended 592x592x336), 3000 steps, 25 Hz Ricker, 128x128 surface receivers
  (stride 4); the shot record is synthesised first by the modeling path.
  value = Mcell-updates/s of the whole shot (updated cells per wavefield step
  x (3 nt - 2) wavefield steps / shot time); also RTM shots/hour and the
  roofline of the dominant stencil kernel (algorithmic bytes / average launch
  duration from HIP events around every launch of the timed region).

Beside the headline, in this order (a leg that would push the whole run past
its time budget is skipped, and the line says so):
  cpu_baseline   the FP32 oracle port of the reference CPU operators timed on
                 this host's cores on the GPU's own (bit-identical) states at
                 9 times of the forward and of the backward pass, integrated
                 over the shot (so the denormal phase is counted where it
                 occurs); and the reference CPU program itself, compiled from
                 its sources, on the 128^3 three-layer modeling case in full.
                 nproc, OMP_NUM_THREADS and the CPU model are reported.
  snapshot       compressed-snapshot RTM of the salt shot (FP16 wavefields,
                 FP16 snapshots every 4 steps; FP32 wavefields with FP16
                 snapshots and with autoencoder-coded snapshots) against the
                 FP32 HBC shot of the headline: shots/hour ratio.
  three_layer_128  the 128^3 three-layer modeling case (1000 steps) on the GPU.
  accuracy       a 128^3 crop of the salt model (border 16), 4000 steps:
                 traces and HBC images of each wavefield format against the
                 FP64 oracle; compressed-snapshot images (FP16 and coded,
                 the codec trained on two other shots' snapshots) against
                 the exact-snapshot (FP32, every step) image.
  precision      the 800x800x400 salt model (extended 880x880x480), 4000 steps:
                 HBC shots with FP16 and BF16 wavefields (and FP32 when the
                 budget allows).
  batched        all 64 windows (uneven apertures) of the 1024x1024x320 model:
                 one at a time vs one batch (FP32), and the batch with FP16
                 wavefields; nt reduced from 2500 (--batch-nt), the GPU's cost
                 per step does not depend on nt.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
       [--no-tune] [--skip LEG,...] [--budget SECONDS]
"""
import argparse
import concurrent.futures as cf
import hashlib
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = ("Mcell-updates/s + RTM shots/hour at image rel-L2≤1e-2, "
          "batched/learned vs plain FP32")
HBM_PEAK_GBS = 8000.0
REFBIN = os.path.join(REPO, "oracle", "_ref", "RTM3D.bin")
T0 = time.perf_counter()


def elapsed():
    return time.perf_counter() - T0


def log(*a):
    print("[bench %6.1fs]" % elapsed(), *a, file=sys.stderr, flush=True)


class Budget:
    """Skips optional legs that would push the run past `total` seconds."""

    def __init__(self, total, skip):
        self.total, self.skip, self.skipped = total, set(skip), []

    def allows(self, leg, estimate_s):
        if leg in self.skip:
            self.skipped.append({"leg": leg, "reason": "--skip"})
            return False
        if elapsed() + estimate_s > self.total:
            log("skipping %s: about %.0f s left, it needs about %.0f s"
                % (leg, self.total - elapsed(), estimate_s))
            self.skipped.append({"leg": leg, "reason": "time budget",
                                 "estimate_s": estimate_s})
            return False
        return True


def rel_l2(a, r):
    a = np.asarray(a, np.float64)
    r = np.asarray(r, np.float64)
    return float(np.linalg.norm(a - r) / np.linalg.norm(r))


def host_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = None
    return {"nproc": os.cpu_count(), "cpus_allowed": affinity,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "cpu_model": model}


def cpu_threads():
    """Host threads for the CPU legs: OMP_NUM_THREADS when set (the host's
    share of cores), else every CPU this process may use."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


# ----------------------------------------------------------------- salt shot --
class Shot:
    """Device state of one RTM shot on the salt model."""

    def __init__(self, hip, host, workloads, nt):
        self.hip = hip
        p = workloads.salt_512_params(nt)
        self.p = p
        t = time.time()
        vel = workloads.salt_model()
        log("velocity model %s built in %.1f s" % (vel.shape, time.time() - t))
        b, order = p["blen"], p["stencil_order"]
        self.b, self.order, self.nt = b, order, nt
        ve_abc = host.extend_velocity(vel, b, "abc")
        self.shape = ve_abc.shape
        self.coefs = host.stencil_coefs(order, p["dx"], p["dy"], p["dz"])
        self.V = hip.DeviceArray.from_host(host.v2dt2(ve_abc, p["dt"]))
        del ve_abc
        self.src_host = host.ricker(nt, p["dt"], p["fpeak"])
        self.g_host = host.taper_table(b, p["taper_factor"])
        self.src = hip.DeviceArray.from_host(self.src_host)
        self.g = hip.DeviceArray.from_host(self.g_host)
        self.geom = hip.RcvGeom(p["receiver_start_x"], p["receiver_start_y"],
                                p["receiver_depth_z"], p["receiver_distance_x"],
                                p["receiver_distance_y"], p["receiver_count_x"],
                                p["receiver_count_y"])
        self.src_ext = (p["source_start_x"] + b, p["source_start_y"] + b, p["source_depth_z"] + b)
        self.shot = hip.make_shot(*self.shape, b, order, nt, nt, self.src_ext, self.geom)
        self.F = hip.FieldPair(self.shape)
        self.R = hip.FieldPair(self.shape)
        self.extra = {}  # 16-bit wavefield buffers per format
        self.rcv = hip.DeviceArray((p["receiver_count_x"], p["receiver_count_y"], nt)).zero()
        self.img = hip.DeviceArray(self.shape).zero()
        self.upb = hip.DeviceArray((nt, self.shape[0], self.shape[1], order // 2))
        # modeling (ABC velocity) synthesises the record; migration uses HBC
        hip.synchronize()
        t = time.perf_counter()
        hip.model_steps(self.shot, self.V, self.coefs, self.g, self.src, self.F, self.rcv, 0, nt)
        hip.synchronize()
        self.modeling_s = time.perf_counter() - t
        self.V.upload(host.v2dt2(host.extend_velocity(vel, b, "hbc"), p["dt"]))
        del vel
        self.U = float((self.shape[0] - order) * (self.shape[1] - order) * (self.shape[2] - order))

    def fields(self, fmt):
        """(shot, source pair, receiver pair) for a wavefield storage format."""
        if fmt == "f32":
            return self.shot, self.F, self.R
        if fmt not in self.extra:
            dt = self.hip.WAVE_FORMATS[fmt][1]
            shot = self.hip.make_shot(*self.shape, self.b, self.order, self.nt, self.nt,
                                      self.src_ext, self.geom, fmt)
            self.extra[fmt] = (shot, self.hip.FieldPair(self.shape, dt),
                               self.hip.FieldPair(self.shape, dt))
        return self.extra[fmt]

    def release(self, fmt):
        if fmt in self.extra:
            for x in self.extra.pop(fmt)[1:]:
                x.free()

    def migrate(self):
        """One FP32 HBC shot (the headline step)."""
        hip = self.hip
        for f in (self.F.a, self.F.b, self.R.a, self.R.b, self.img):
            f.zero()
        hip.forward_steps(self.shot, True, self.V, self.coefs, self.g, self.src, self.F,
                          self.upb, 0, self.nt)
        # mode 1: fused receiver step (the one-pass source + receiver mode 2
        # is bit-identical but, with its 4-row tiles, not faster here)
        hip.backward_steps(self.shot, True, self.V, self.coefs, self.g, self.F, self.R,
                           self.upb, self.rcv, self.img, self.nt - 1, self.nt, 1)

    def free(self):
        for fmt in list(self.extra):
            self.release(fmt)
        for x in (self.V, self.src, self.g, self.F, self.R, self.rcv, self.img, self.upb):
            x.free()


def read_traffic(cls, tile):
    """HBM bytes per launch of stencil class `cls` with tile shape `tile`
    ("TZxBYxRY"), from a committed rocprofv3 PMC summary
    ("traffic_bytes_per_launch[class C, tile T]: bytes"), or None."""
    pdir = os.path.join(REPO, "profiles")
    pat = re.compile(r"traffic_bytes_per_launch\[class %d, tile %s\]:\s*([0-9.eE+]+)"
                     % (cls, re.escape(tile)))
    for name in sorted(os.listdir(pdir)) if os.path.isdir(pdir) else []:
        if name.endswith(".md"):
            m = pat.search(open(os.path.join(pdir, name)).read())
            if m:
                return float(m.group(1)), name
    return None, None


# -------------------------------------------------------------- CPU leg -------
def _oracle():
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import pyoracle
    pyoracle.lib()
    return pyoracle


def cpu_salt_phases(hip, sh, samples=(0, 60, 150, 300, 600, 1000, 1600, 2300, 2990)):
    """Per-step cost of the FP32 oracle port of the reference CPU operators on
    the salt shot, measured from the GPU's states (bit-identical to the
    reference's) at sample times of the forward pass (taper x2, step,
    source, upper-border save) and of the backward pass (reconstruction step,
    upper-border restore, receiver tapers x2, receiver step, receiver
    injection, imaging), two steps each, then integrated over the shot with
    piecewise-linear interpolation.  The GPU replays the shot in chunks to
    reach each sample."""
    po = _oracle()
    nt, b, order = sh.nt, sh.b, sh.order
    v2 = sh.V.download()
    rcv = sh.rcv.download()
    geom = po.RcvGeom(sh.geom.start_x, sh.geom.start_y, sh.geom.depth_z, sh.geom.dist_x,
                      sh.geom.dist_y, sh.geom.count_x, sh.geom.count_y)
    coefs, g, src = sh.coefs, sh.g_host, sh.src_host
    upb1 = np.zeros((1, sh.shape[0], sh.shape[1], order // 2), np.float32)  # one layer
    samples = sorted(t for t in samples if 0 <= t <= nt - 3)
    F, R = sh.F, sh.R
    for f in (F.a, F.b, R.a, R.b, sh.img):
        f.zero()
    fwd, done = [], 0
    for t in samples:
        if t > done:
            hip.forward_steps(sh.shot, True, sh.V, sh.coefs, sh.g, sh.src, F, sh.upb, done,
                              t - done)
            done = t
        P, PP = F.P.download(), F.PP.download()
        c0 = time.perf_counter()
        for it in (t, t + 1):
            P, PP = PP, P
            po.taper(P, g, b, True)
            po.taper(PP, g, b, True)
            po.step(P, PP, v2, coefs, order)
            po.apply_source(PP, *sh.src_ext, float(src[it]))
            po.upb(PP, upb1, b, order, 0, False)
        fwd.append((t, (time.perf_counter() - c0) / 2))
        log("cpu port forward step at t=%d: %.3f s" % (t, fwd[-1][1]))
    hip.forward_steps(sh.shot, True, sh.V, sh.coefs, sh.g, sh.src, F, sh.upb, done, nt - done)
    bwd, hi = [], nt - 1
    img = np.zeros(sh.shape, np.float32)
    for t in sorted(samples, reverse=True):
        t = min(t, nt - 3)
        if hi > t:
            hip.backward_steps(sh.shot, True, sh.V, sh.coefs, sh.g, F, R, sh.upb, sh.rcv,
                               sh.img, hi, hi - t, 1)
            hi = t
        Ps, PPs, Pr, PPr = F.P.download(), F.PP.download(), R.P.download(), R.PP.download()
        c0 = time.perf_counter()
        for it in (t, t - 1):
            if it < 0:
                break
            po.step(Ps, PPs, v2, coefs, order)
            po.upb(PPs, upb1, b, order, 0, True)
            Ps, PPs = PPs, Ps
            po.taper(Pr, g, b, True)
            po.taper(PPr, g, b, True)
            po.step(Pr, PPr, v2, coefs, order)
            po.receivers(PPr, rcv, b, geom, nt, it, True)
            po.image(img, Ps, PPr, b)
            Pr, PPr = PPr, Pr
        bwd.append((t, (time.perf_counter() - c0) / (2 if t >= 1 else 1)))
        log("cpu port backward step at t=%d: %.3f s" % (t, bwd[-1][1]))
    hip.synchronize()
    ts = np.arange(nt)
    f_cost = np.interp(ts, [t for t, _ in fwd], [c for _, c in fwd])
    b_cost = np.interp(ts, sorted(t for t, _ in bwd), [c for _, c in sorted(bwd)])
    shot_s = float(f_cost.sum() + b_cost.sum())
    return {
        "shot_s_extrapolated": round(shot_s, 1),
        "forward_step_s": {str(t): round(c, 4) for t, c in fwd},
        "backward_step_s": {str(t): round(c, 4) for t, c in sorted(bwd)},
        "mcell_updates_per_s": round(sh.U * (3 * nt - 2) / shot_s / 1e6, 2),
    }


def cpu_three_layer_reference(workloads, threads, timeout_s, gpu_record=None):
    """The reference CPU program, compiled from its sources, on the 128^3
    three-layer modeling case in full (1000 steps), timed with a microsecond
    wall clock around the process and by its own whole-millisecond report."""
    if not os.path.exists(REFBIN):
        return {"error": "reference program not built (oracle/_ref/RTM3D.bin)"}
    wd = tempfile.mkdtemp(prefix="rtmcpu_")
    try:
        for d in ("vel", "out", "tmp", "dat"):
            os.makedirs(os.path.join(wd, d))
        workloads.three_layer(128, 128, 128, 40, 80).tofile(os.path.join(wd, "vel", "model.bin"))
        p = workloads.three_layer_128_params()
        p["nthreads"] = threads
        with open(os.path.join(wd, "in.json"), "w") as f:
            json.dump(p, f)
        t = time.perf_counter()
        r = subprocess.run([REFBIN, "in.json"], cwd=wd, capture_output=True, text=True,
                           timeout=timeout_s, env=dict(os.environ, OMP_NUM_THREADS=str(threads)))
        wall = time.perf_counter() - t
        if r.returncode != 0:
            return {"error": "reference run failed: " + r.stdout[-300:]}
        m = re.search(r'"modeling_time":\s*([0-9.eE+-]+)', r.stdout)
        name = "RTMSEISMOGRAM_S64x64x2_NXNY_64x64_NTSTART0_NTSTEP1000.seism"
        raw = open(os.path.join(wd, "dat", name), "rb").read()
        out = {"wall_s": round(wall, 3), "reference_modeling_time_s": float(m.group(1)) if m
               else None, "threads": threads,
               "mcell_updates_per_s": round(6229504.0 * 1000 / wall / 1e6, 2),
               "record_sha256": hashlib.sha256(raw).hexdigest()}
        if gpu_record is not None:
            out["gpu_record_bit_identical"] = bool(raw == gpu_record)
        return out
    except subprocess.TimeoutExpired:
        return {"error": "reference run exceeded %.0f s" % timeout_s, "threads": threads}
    finally:
        shutil.rmtree(wd, ignore_errors=True)


# ------------------------------------------------------- GPU modeling case ----
def three_layer_gpu(hip, host, workloads, reps=3):
    """The 128^3 three-layer modeling case (1000 steps) on the GPU: the
    record and the Mcell-updates/s of the best of `reps` runs."""
    p = workloads.three_layer_128_params()
    b, order, nt = p["blen"], p["stencil_order"], p["nt"]
    vel = workloads.three_layer(128, 128, 128, 40, 80)
    ve = host.extend_velocity(vel, b, "abc")
    V = hip.DeviceArray.from_host(host.v2dt2(ve, p["dt"]))
    coefs = host.stencil_coefs(order, p["dx"], p["dy"], p["dz"])
    src = hip.DeviceArray.from_host(host.ricker(nt, p["dt"], p["fpeak"]))
    g = hip.DeviceArray.from_host(host.taper_table(b, p["taper_factor"]))
    geom = hip.RcvGeom(0, 0, 0, 2, 2, 64, 64)
    shot = hip.make_shot(*ve.shape, b, order, nt, nt, (64 + b, 64 + b, 2 + b), geom)
    F = hip.FieldPair(ve.shape)
    rcv = hip.DeviceArray((64, 64, nt))
    best = None
    for _ in range(reps + 1):  # the first run warms the kernels
        for f in (F.a, F.b, rcv):
            f.zero()
        F.p.value, F.pp.value = F.a.ptr.value, F.b.ptr.value
        hip.synchronize()
        t = time.perf_counter()
        hip.model_steps(shot, V, coefs, g, src, F, rcv, 0, nt)
        hip.synchronize()
        dt = time.perf_counter() - t
        best = dt if best is None else min(best, dt)
    rec = rcv.download()
    for x in (V, src, g, F, rcv):
        x.free()
    U = 6229504.0
    return rec.tobytes(), {"workload": "three_layer_128_nt1000", "grid": [128, 128, 128],
                           "extended_grid": list(ve.shape), "border": b, "nt": nt,
                           "run_s": round(best, 4),
                           "mcell_updates_per_s": round(U * nt / best / 1e6, 1),
                           "note": "3 grids of 28 MB: resident in the 256 MB Infinity Cache, "
                                   "not an HBM roofline case"}


# --------------------------------------------------- frequency-domain leg ---
def freq_leg(hip, host, sh, nsteps=60, nwstep=4):
    """Frequency-domain migration steps (rtmhip_freq_steps) on the salt grid:
    per time step both fields stepped, taper3D of the four grids, source,
    receivers (a synthetic record) and the contributions and image of a group
    of nwstep frequencies -- with the tapers on the new fields' output (the
    default) and as the reference's in-place passes (mode 0); the images of
    the two runs are compared bit for bit."""
    p, nt = sh.p, sh.nt
    w2, kre, kim = host.freq_tables(nt, np.float32(p["dt"]), np.float32(2.0), np.float32(40.0))
    nw = len(w2)
    inner = tuple(n - 2 * sh.b for n in sh.shape)
    acc = [hip.DeviceArray((nwstep,) + inner) for _ in range(4)]
    dk = [hip.DeviceArray.from_host(np.ascontiguousarray(a)) for a in (kre, kim, w2)]
    rng = np.random.default_rng(3)
    rec = hip.DeviceArray.from_host(
        (rng.standard_normal((p["receiver_count_x"], p["receiver_count_y"], nt)) * 1e-3)
        .astype(np.float32))
    img = hip.DeviceArray(sh.shape)
    out, ref = {}, None
    try:
        for name, mode in (("output_taper", -1), ("in_place_tapers", 0)):
            hip.set_taper_fusion("f32", mode)
            for f in (sh.F.a, sh.F.b, sh.R.a, sh.R.b, img, *acc):
                f.zero()
            for fp in (sh.F, sh.R):
                fp.p.value, fp.pp.value = fp.a.ptr.value, fp.b.ptr.value
            hip.synchronize()
            t = time.perf_counter()
            hip.freq_steps(sh.shot, sh.V, sh.coefs, sh.g, sh.src, sh.F, sh.R, rec, img, acc,
                           dk[0], dk[1], dk[2], nwstep, 0, nw, nt - 1, nsteps)
            hip.synchronize()
            dt = time.perf_counter() - t
            im = img.download()
            same = ref is None or np.array_equal(im.view(np.uint32), ref.view(np.uint32))
            ref = im if ref is None else ref
            out[name] = {"ms_per_step": round(dt / nsteps * 1e3, 3), "images_bit_identical": same}
        out["speedup"] = round(out["in_place_tapers"]["ms_per_step"]
                               / out["output_taper"]["ms_per_step"], 3)
        out.update(workload="salt 512x512x256 + HBC 40, %d time steps of both fields" % nsteps,
                   frequencies_per_group=nwstep, frequencies_total=nw,
                   note="the reference runs nw/nwstep groups of nt steps each")
        log("frequency-domain steps: %.3f ms per step (%.3f with in-place tapers)"
            % (out["output_taper"]["ms_per_step"], out["in_place_tapers"]["ms_per_step"]))
    finally:
        hip.set_taper_fusion("f32", -1)
        for x in acc + dk + [rec, img]:
            x.free()
    return out


# ------------------------------------------------------------ snapshot leg ---
def snapshot_leg(hip, sh, f32_shot_s, k=4):
    """Compressed-snapshot RTM of the salt shot against the FP32 HBC shot."""
    out = {"snapshot_interval": k, "imaging": "every %d-th step, weighted by %d" % (k, k)}
    saved = [hip.get_tiling(c) for c in range(3)]
    for c in range(3):
        hip.set_tiling(c, 0)  # 16-bit storage runs its own default shapes
    for wfmt, sfmt in (("f16", "f16"), ("f32", "f16")):
        if wfmt == "f32":
            for c, tl in enumerate(saved):
                hip.set_tiling(c, *tl) if tl[0] else hip.set_tiling(c, 0)
        shot, F, R = sh.fields(wfmt)
        snaps = hip.snap_buffer(shot, sfmt, k)
        for f in (F.a, F.b, R.a, R.b, sh.img):
            f.zero()
        F.p.value, F.pp.value = F.a.ptr.value, F.b.ptr.value
        R.p.value, R.pp.value = R.a.ptr.value, R.b.ptr.value
        hip.synchronize()
        t = time.perf_counter()
        hip.snap_forward_steps(shot, sh.V, sh.coefs, sh.g, sh.src, F, snaps, sfmt, k, 0, sh.nt)
        hip.snap_backward_steps(shot, sh.V, sh.coefs, sh.g, R, snaps, sfmt, k, sh.rcv, sh.img,
                                sh.nt - 1, sh.nt)
        hip.synchronize()
        s = time.perf_counter() - t
        img = sh.img.download()
        if not np.isfinite(img).all() or not np.abs(img).max() > 0:
            raise RuntimeError("invalid snapshot-RTM image")
        del img
        key = "%s_wavefields_%s_snapshots" % (wfmt, sfmt)
        out[key] = {"shot_s": round(s, 3), "shots_per_hour": round(3600.0 / s, 1),
                    "speedup_vs_f32_hbc": round(f32_shot_s / s, 3),
                    "snapshot_GB": round(snaps.nbytes / 1e9, 1)}
        log("snapshot RTM %s: %.3f s (%.2fx the FP32 HBC shot)" % (key, s, f32_shot_s / s))
        snaps.free()
        sh.release(wfmt)
    for c, tl in enumerate(saved):
        hip.set_tiling(c, *tl) if tl[0] else hip.set_tiling(c, 0)
    # transform-coded snapshots (keep 1/32 of the coefficients per snapshot)
    for wfmt in ("f16", "f32"):
        if wfmt == "f16":
            for c in range(3):
                hip.set_tiling(c, 0)
        else:
            for c, tl in enumerate(saved):
                hip.set_tiling(c, *tl) if tl[0] else hip.set_tiling(c, 0)
        shot, F, R = sh.fields(wfmt)
        st = hip.DctSnapshots(shot, k, 32)
        for f in (F.a, F.b, R.a, R.b, sh.img):
            f.zero()
        F.p.value, F.pp.value = F.a.ptr.value, F.b.ptr.value
        R.p.value, R.pp.value = R.a.ptr.value, R.b.ptr.value
        hip.synchronize()
        t = time.perf_counter()
        st.forward_steps(sh.V, sh.coefs, sh.g, sh.src, F, 0, sh.nt)
        st.backward_steps(sh.V, sh.coefs, sh.g, R, sh.rcv, sh.img, sh.nt - 1, sh.nt)
        hip.synchronize()
        s = time.perf_counter() - t
        img = sh.img.download()
        if not np.isfinite(img).all() or not np.abs(img).max() > 0:
            raise RuntimeError("invalid DCT-snapshot RTM image")
        del img
        key = "%s_wavefields_dct_snapshots" % wfmt
        out[key] = {"shot_s": round(s, 3), "shots_per_hour": round(3600.0 / s, 1),
                    "speedup_vs_f32_hbc": round(f32_shot_s / s, 3),
                    "snapshot_GB": round(st.nbytes() / 1e9, 2),
                    "coding": "3D DCT-II on 8x8x32 bricks, at most 1/32 of the coefficients "
                              "per snapshot (one threshold per snapshot), FP16 values"}
        log("snapshot RTM %s: %.3f s (%.2fx the FP32 HBC shot), %.2f GB"
            % (key, s, f32_shot_s / s, st.nbytes() / 1e9))
        st.free()
        sh.release(wfmt)
    for c, tl in enumerate(saved):
        hip.set_tiling(c, *tl) if tl[0] else hip.set_tiling(c, 0)
    # coded snapshots (FP32 wavefields): the work per snapshot does not depend
    # on the weights, so a seeded network times it; the accuracy leg images
    # with a trained one
    from rtm3d_demo_amd import ae
    codec = ae.Codec(max_bricks=1024, seed=1)
    shot, F, R = sh.fields("f32")
    cs = ae.CodedSnapshots(shot, k, codec)
    for f in (F.a, F.b, R.a, R.b, sh.img):
        f.zero()
    F.p.value, F.pp.value = F.a.ptr.value, F.b.ptr.value
    R.p.value, R.pp.value = R.a.ptr.value, R.b.ptr.value
    hip.synchronize()
    t = time.perf_counter()
    cs.forward_steps(sh.V, sh.coefs, sh.g, sh.src, F, 0, sh.nt)
    cs.backward_steps(sh.V, sh.coefs, sh.g, R, sh.rcv, sh.img, sh.nt - 1, sh.nt)
    hip.synchronize()
    s = time.perf_counter() - t
    img = sh.img.download()
    if not np.isfinite(img).all() or not np.abs(img).max() > 0:
        raise RuntimeError("invalid coded-snapshot RTM image")
    del img
    out["f32_wavefields_coded_snapshots"] = {
        "shot_s": round(s, 3), "shots_per_hour": round(3600.0 / s, 1),
        "speedup_vs_f32_hbc": round(f32_shot_s / s, 3),
        "snapshot_GB": round(cs.nbytes() / 1e9, 2),
        "bricks_per_snapshot": cs.nb,
        "codec": "32^3 bricks -> 2048 FP16 codes + 1 FP32 scale (matrix-core encode/decode)"}
    log("snapshot RTM f32_wavefields_coded_snapshots: %.3f s (%.2fx the FP32 HBC shot), "
        "%.2f GB of codes" % (s, f32_shot_s / s, cs.nbytes() / 1e9))
    cs.free()
    codec.destroy()
    return out


# ----------------------------------------------------------- accuracy leg ----
def accuracy_crop(hip, host, workloads, formats=("f32", "f16", "bf16"), crop=(128, 128, 128),
                  nt=4000, snap=(("f16", "f16", 4), ("f32", "f16", 4), ("f16", "f16", 8))):
    """Device results vs the FP64 oracle on a crop of the salt model (border
    16): relative L2 of the receiver traces and of the migrated HBC image for
    each wavefield format (each format models its own record and migrates
    it); and snapshot-RTM images (wavefield format, snapshot format, interval)
    vs the exact-snapshot image (FP32 wavefields, FP32 snapshots of every
    step) of the FP32 record."""
    po = _oracle()
    cx, cy, cz = crop
    full = workloads.salt_model()
    x0, y0 = 256 - cx // 2, 256 - cy // 2
    vel = np.ascontiguousarray(full[x0:x0 + cx, y0:y0 + cy, :cz])
    del full
    d, b, order = 12.5, 16, 8
    nr = (cx // 4, cy // 4)
    p = workloads.params("crop", cx, cy, cz, d, b, nt, workloads.cfl_dt(4500.0, d), 25.0,
                         (cx // 2, cy // 2, 2), (0, 0, 0, 4, 4) + nr)
    geom = hip.RcvGeom(0, 0, 0, 4, 4, *nr)
    coefs = host.stencil_coefs(order, d, d, d)
    src = host.ricker(nt, p["dt"], p["fpeak"])
    g = host.taper_table(b, p["taper_factor"])
    ve = host.extend_velocity(vel, b, "abc")
    vpe = host.extend_velocity(vel, b, "hbc")
    shape = ve.shape
    V = hip.DeviceArray.from_host(host.v2dt2(ve, p["dt"]))
    Vh = hip.DeviceArray.from_host(host.v2dt2(vpe, p["dt"]))
    dsrc, dg = hip.DeviceArray.from_host(src), hip.DeviceArray.from_host(g)
    fcoefs = host.stencil_coefs(2, d, d, d)
    out_img = hip.DeviceArray(crop)

    def finish(img):
        hip.image_finish(img, out_img, b, fcoefs)
        return out_img.download()

    results, rcv32 = {}, None
    for fmt in formats:
        wdt = hip.WAVE_FORMATS[fmt][1]
        shot = hip.make_shot(*shape, b, order, nt, nt, (cx // 2 + b, cy // 2 + b, 2 + b), geom,
                             fmt)
        F = hip.FieldPair(shape, wdt)
        rcv = hip.DeviceArray(nr + (nt,)).zero()
        hip.model_steps(shot, V, coefs, dg, dsrc, F, rcv, 0, nt)
        rec = rcv.download()
        S, R = hip.FieldPair(shape, wdt), hip.FieldPair(shape, wdt)
        upb = hip.DeviceArray((nt, shape[0], shape[1], 4), wdt)
        img = hip.DeviceArray(shape).zero()
        hip.forward_steps(shot, True, Vh, coefs, dg, dsrc, S, upb, 0, nt)
        hip.backward_steps(shot, True, Vh, coefs, dg, S, R, upb, rcv, img, nt - 1, nt, True)
        results[fmt] = (rec, finish(img))
        for x in (F, S, R, upb, img):
            x.free()
        if fmt == "f32":
            rcv32 = rcv
        else:
            rcv.free()
    # snapshot RTM of the FP32 record: exact (FP32, every step) and compressed
    snap_imgs = {}
    for wfmt, sfmt, k in (("f32", "f32", 1),) + tuple(snap):
        wdt = hip.WAVE_FORMATS[wfmt][1]
        shot = hip.make_shot(*shape, b, order, nt, nt, (cx // 2 + b, cy // 2 + b, 2 + b), geom,
                             wfmt)
        S, R = hip.FieldPair(shape, wdt), hip.FieldPair(shape, wdt)
        snaps = hip.snap_buffer(shot, sfmt, k)
        img = hip.DeviceArray(shape).zero()
        hip.snap_forward_steps(shot, Vh, coefs, dg, dsrc, S, snaps, sfmt, k, 0, nt)
        hip.snap_backward_steps(shot, Vh, coefs, dg, R, snaps, sfmt, k, rcv32, img, nt - 1, nt)
        snap_imgs[(wfmt, sfmt, k)] = finish(img)
        for x in (S, R, snaps, img):
            x.free()
    # transform-coded snapshots every 4th step, keeping 1/32 and 1/64 of the
    # coefficients
    for wfmt, div in (("f32", 32), ("f16", 32), ("f32", 64)):
        wdt = hip.WAVE_FORMATS[wfmt][1]
        shot = hip.make_shot(*shape, b, order, nt, nt, (cx // 2 + b, cy // 2 + b, 2 + b), geom,
                             wfmt)
        S, R = hip.FieldPair(shape, wdt), hip.FieldPair(shape, wdt)
        for f in (S.a, S.b, R.a, R.b):
            f.zero()
        st = hip.DctSnapshots(shot, 4, div)
        img = hip.DeviceArray(shape).zero()
        st.forward_steps(Vh, coefs, dg, dsrc, S, 0, nt)
        st.backward_steps(Vh, coefs, dg, R, rcv32, img, nt - 1, nt)
        snap_imgs[(wfmt, "dct%d" % div, 4)] = finish(img)
        for x in (S, R, img):
            x.free()
        st.free()
    # coded snapshots (every 4th step) with a codec trained on the snapshots of
    # two other shots of the crop (steps after nt/3, every 50th); the same
    # snapshots train the learned basis of the transform coder
    from rtm3d_demo_amd import ae
    from rtm3d_demo_amd.basis import BasisLearner
    learner = BasisLearner()
    t = time.perf_counter()
    lz, kt = shape[2] - 2 * b, 50
    nb = ae.region_bricks((0, 0, 0, shape[0], shape[1], lz))
    slots = [j for j in range((nt - 1) // kt + 1) if nt - 1 - j * kt >= nt // 3]
    train_src = ((cx // 4, 3 * cy // 4), (3 * cx // 4, cy // 4))
    pool = hip.DeviceArray((len(train_src) * len(slots) * nb, ae.BRICK), np.uint16)
    scl = hip.DeviceArray((nb,))
    row = 0
    for sx, sy in train_src:
        tshot = hip.make_shot(*shape, b, order, nt, nt, (sx + b, sy + b, 2 + b), geom)
        S = hip.FieldPair(shape)
        snaps = hip.snap_buffer(tshot, "f32", kt)
        hip.snap_forward_steps(tshot, Vh, coefs, dg, dsrc, S, snaps, "f32", kt, 0, nt)
        row = ae.collect_bricks(snaps, slots, pool, row, scl)
        for j in slots:
            learner.add(snaps.slot(j).download(), b)
        S.free()
        snaps.free()
    codec = ae.Codec(max_bricks=256, seed=3)
    losses = codec.train(pool, row, iters=2000, batch=64, lr=5e-3, seed=5)
    pool.free()
    scl.free()
    shot = hip.make_shot(*shape, b, order, nt, nt, (cx // 2 + b, cy // 2 + b, 2 + b), geom)
    S, R = hip.FieldPair(shape), hip.FieldPair(shape)
    img = hip.DeviceArray(shape).zero()
    cs = ae.CodedSnapshots(shot, 4, codec)
    cs.forward_steps(Vh, coefs, dg, dsrc, S, 0, nt)
    cs.backward_steps(Vh, coefs, dg, R, rcv32, img, nt - 1, nt)
    snap_imgs[("f32", "coded", 4)] = finish(img)
    coded = {"training": "%d bricks from 2 other shots, 2000 Adam steps of 64 bricks, lr 5e-3"
                         % row,
             "train_mse_first_20": float(losses[:20].mean()),
             "train_mse_last_50": float(losses[-50:].mean()),
             "bytes_per_snapshot": cs.nbytes() // cs.nsnap,
             "f32_snapshot_bytes": shape[0] * shape[1] * lz * 4,
             "leg_s": round(time.perf_counter() - t, 1)}
    for x in (S, R, img):
        x.free()
    cs.free()
    codec.destroy()
    # transform-coded snapshots with the learned basis
    basis = learner.basis()
    for wfmt, div in (("f32", 32), ("f16", 32), ("f32", 64)):
        wdt = hip.WAVE_FORMATS[wfmt][1]
        shot = hip.make_shot(*shape, b, order, nt, nt, (cx // 2 + b, cy // 2 + b, 2 + b), geom,
                             wfmt)
        S, R = hip.FieldPair(shape, wdt), hip.FieldPair(shape, wdt)
        for f in (S.a, S.b, R.a, R.b):
            f.zero()
        st = hip.DctSnapshots(shot, 4, div, basis)
        img = hip.DeviceArray(shape).zero()
        st.forward_steps(Vh, coefs, dg, dsrc, S, 0, nt)
        st.backward_steps(Vh, coefs, dg, R, rcv32, img, nt - 1, nt)
        snap_imgs[(wfmt, "dct%d_learned_basis" % div, 4)] = finish(img)
        for x in (S, R, img):
            x.free()
        st.free()
    for x in (V, Vh, dsrc, dg, rcv32, out_img):
        x.free()
    t = time.perf_counter()
    rec64 = po.pipeline_modeling(p, vel, np.float64)[0, 0]
    q = dict(p, modeling=False, migration=True)
    img64 = po.pipeline_migration(q, vpe, rec64[None, None], np.float64)
    log("accuracy: FP64 oracle modeling + HBC migration in %.1f s" % (time.perf_counter() - t))
    exact = snap_imgs[("f32", "f32", 1)]
    acc = {fmt: {"trace_rel_l2": rel_l2(rec, rec64), "image_rel_l2": rel_l2(im, img64)}
           for fmt, (rec, im) in results.items()}
    snap_acc = {"exact_snapshot_vs_fp64_hbc_image_rel_l2": rel_l2(exact, img64),
                "fp32_hbc_vs_exact_snapshot_image_rel_l2": rel_l2(results["f32"][1], exact)}
    for (wfmt, sfmt, k), im in snap_imgs.items():
        if k == 1 and sfmt == "f32":
            continue
        snap_acc["%s_wavefields_%s_snapshots_k%d_image_rel_l2_vs_exact_snapshot"
                 % (wfmt, sfmt, k)] = rel_l2(im, exact)
    snap_acc["coded_snapshot_codec"] = coded
    snap_acc["learned_basis_training"] = ("principal directions of the z, x and y lines of "
                                          "the 8x8x32 bricks of %d snapshots of the 2 "
                                          "training shots" % (2 * len(slots)))
    case = "%dx%dx%d crop of the salt model, border 16, %d steps, %dx%d receivers" % (
        cx, cy, cz, nt, nr[0], nr[1])
    return acc, snap_acc, case


# -------------------------------------------------- 880 x 880 x 480 shots ----
def precision_leg(hip, host, workloads, budget, nt=4000, f32_rate=None):
    """HBC shots on the 800x800x400 salt model (extended 880x880x480) with
    FP16 and BF16 wavefield storage (FP32 arithmetic), and FP32 when the time
    budget allows; the record comes from the FP32 modeling path."""
    d, b, order = 12.5, 40, 8
    t = time.time()
    vel = workloads.salt_model(800, 800, 400)
    log("800x800x400 model built in %.1f s" % (time.time() - t))
    p = workloads.params("salt800", 800, 800, 400, d, b, nt, workloads.cfl_dt(4500.0, d), 25.0,
                         (400, 400, 2), (0, 0, 0, 4, 4, 200, 200))
    coefs = host.stencil_coefs(order, d, d, d)
    V = hip.DeviceArray.from_host(host.v2dt2(host.extend_velocity(vel, b, "abc"), p["dt"]))
    shape = V.shape
    src = hip.DeviceArray.from_host(host.ricker(nt, p["dt"], p["fpeak"]))
    g = hip.DeviceArray.from_host(host.taper_table(b, p["taper_factor"]))
    geom = hip.RcvGeom(0, 0, 0, 4, 4, 200, 200)
    sxe = (400 + b, 400 + b, 2 + b)
    rcv = hip.DeviceArray((200, 200, nt)).zero()
    F = hip.FieldPair(shape)
    shot32 = hip.make_shot(*shape, b, order, nt, nt, sxe, geom)
    hip.synchronize()
    t = time.perf_counter()
    hip.model_steps(shot32, V, coefs, g, src, F, rcv, 0, nt)
    hip.synchronize()
    model_s = time.perf_counter() - t
    F.free()
    V.upload(host.v2dt2(host.extend_velocity(vel, b, "hbc"), p["dt"]))
    del vel
    U = float((shape[0] - 8) * (shape[1] - 8) * (shape[2] - 8))
    updates = U * (3 * nt - 2)
    out = {"workload": "hbc_rtm_salt800_nt%d" % nt, "grid": [800, 800, 400],
           "extended_grid": list(shape), "dx_m": d, "border": b, "nt": nt,
           "modeling_s": round(model_s, 3),
           "snapshots": "the HBC shots reconstruct the source field (FP16 snapshots of every "
                        "4th step over the imaging range would take 1000 x 880 x 880 x 400 x "
                        "2 B = 620 GB, more than the GPU's 288 GB); snapshots every 4th step "
                        "run transform-coded (f16_wavefields_dct_snapshots_k4)"}
    saved = [hip.get_tiling(c) for c in range(3)]
    img = hip.DeviceArray(shape)
    for fmt in ("f16", "bf16", "f32"):
        if fmt == "f32" and not budget.allows("precision_f32", 30.0):
            break
        if fmt == "f32":
            for c, tl in enumerate(saved):
                hip.set_tiling(c, *tl) if tl[0] else hip.set_tiling(c, 0)
        else:
            for c in range(3):
                hip.set_tiling(c, 0)
        wdt = hip.WAVE_FORMATS[fmt][1]
        shot = hip.make_shot(*shape, b, order, nt, nt, sxe, geom, fmt)
        S, R = hip.FieldPair(shape, wdt), hip.FieldPair(shape, wdt)
        upb = hip.DeviceArray((nt, shape[0], shape[1], order // 2), wdt)
        img.zero()
        hip.synchronize()
        hip.timing_enable(True)
        t = time.perf_counter()
        hip.forward_steps(shot, True, V, coefs, g, src, S, upb, 0, nt)
        hip.backward_steps(shot, True, V, coefs, g, S, R, upb, rcv, img, nt - 1, nt, True)
        hip.synchronize()
        s = time.perf_counter() - t
        cls = {c: hip.timing_read(c) for c in range(4)}
        hip.timing_enable(False)
        for x in (S, R, upb):
            x.free()
        im = img.download()
        if not np.isfinite(im).all():
            raise RuntimeError("non-finite image")
        del im
        bpc = 4.0 if fmt == "f32" else 2.0
        r = {"shot_s": round(s, 3), "mcell_updates_per_s": round(updates / s / 1e6, 1),
             "shots_per_hour": round(3600.0 / s, 1)}
        n0, m0 = cls[0]
        if n0:
            a0 = m0 / n0 * 1e-3
            r["plain_step_avg_ms"] = round(a0 * 1e3, 4)
            r["plain_step_achieved_GBs"] = round((4.0 + 3 * bpc) * U / a0 / 1e9, 1)
        out[fmt] = r
        log("precision leg %s: %.3f s" % (fmt, s))
        if fmt == "f16":
            # snapshot RTM: FP16 wavefields, transform-coded snapshots of
            # every 4th step (keep 1/32), instead of HBC reconstruction
            S, R = hip.FieldPair(shape, wdt), hip.FieldPair(shape, wdt)
            for f in (S.a, S.b, R.a, R.b):
                f.zero()
            st = hip.DctSnapshots(shot, 4, 32)
            img.zero()
            hip.synchronize()
            t = time.perf_counter()
            st.forward_steps(V, coefs, g, src, S, 0, nt)
            st.backward_steps(V, coefs, g, R, rcv, img, nt - 1, nt)
            hip.synchronize()
            s = time.perf_counter() - t
            out["f16_wavefields_dct_snapshots_k4"] = {
                "shot_s": round(s, 3), "shots_per_hour": round(3600.0 / s, 1),
                "snapshot_GB": round(st.nbytes() / 1e9, 2)}
            log("precision leg f16 + DCT snapshots: %.3f s, %.1f GB" % (s, st.nbytes() / 1e9))
            for x in (S, R, st):
                x.free()
    for c, tl in enumerate(saved):
        hip.set_tiling(c, *tl) if tl[0] else hip.set_tiling(c, 0)
    if "f32" in out:
        for fmt in ("f16", "bf16", "f16_wavefields_dct_snapshots_k4"):
            if fmt in out:
                out[fmt]["speedup_vs_f32"] = round(out["f32"]["shot_s"] / out[fmt]["shot_s"], 3)
    for x in (V, src, g, rcv, img):
        x.free()
    return out


# ------------------------------------------------------------ batched leg ----
def batch_leg(hip, host, workloads, nt, nshots=64):
    """All `nshots` windows of the 1024x1024x320 model: one at a time through
    the per-shot loops (plain FP32 single-shot GPU RTM) against one batch
    (rtmhip_batch_*: per-step launches of all shots, graph replays), images
    compared bit for bit; then the batch with FP16 wavefields."""
    d, b, order = 12.5, 40, 8
    dt = workloads.cfl_dt(4500.0, d)
    wins = workloads.salt_1024_shots(64)[:nshots]
    t = time.time()
    model = workloads.salt_1024_model()
    coefs = host.stencil_coefs(order, d, d, d)
    g = hip.DeviceArray.from_host(host.taper_table(b, np.float32(0.9)))
    src = hip.DeviceArray.from_host(host.ricker(nt, dt, 25.0))

    def prep(w):
        x0, y0, nxy, nz, _ = w
        vel = np.ascontiguousarray(model[x0:x0 + nxy, y0:y0 + nxy, :nz])
        return (host.v2dt2(host.extend_velocity(vel, b, "abc"), dt),
                host.v2dt2(host.extend_velocity(vel, b, "hbc"), dt))

    members, grids, U = [], [], 0.0
    with cf.ThreadPoolExecutor(max_workers=min(16, cpu_threads())) as ex:
        for k0 in range(0, nshots, 16):
            for w, (va, vh) in zip(wins[k0:k0 + 16], ex.map(prep, wins[k0:k0 + 16])):
                nxy, nz, s = w[2], w[3], w[4]
                shape = va.shape
                geom = hip.RcvGeom(0, 0, 0, 4, 4, nxy // 4, nxy // 4)
                shot = hip.make_shot(*shape, b, order, nt, nt, (s[0] + b, s[1] + b, s[2] + b),
                                     geom)
                m = dict(shot=shot, V=hip.DeviceArray.from_host(va), ricker=src,
                         S=hip.FieldPair(shape), R=hip.FieldPair(shape),
                         rcv=hip.DeviceArray((nxy // 4, nxy // 4, nt)).zero(),
                         img=hip.DeviceArray(shape).zero(),
                         upb=hip.DeviceArray((nt, shape[0], shape[1], order // 2)))
                hip.model_steps(shot, m["V"], coefs, g, src, m["S"], m["rcv"], 0, nt)
                m["V"].upload(vh)
                members.append(m)
                grids.append(list(shape))
                U += float((shape[0] - 8) * (shape[1] - 8) * (shape[2] - 8))
    del model
    hip.synchronize()
    log("batch: %d windows built and modelled in %.1f s" % (nshots, time.time() - t))

    def reset(mem):
        for m in mem:
            for f in (m["S"], m["R"]):
                f.a.zero()
                f.b.zero()
                f.p.value, f.pp.value = f.a.ptr.value, f.b.ptr.value
            m["img"].zero()

    reset(members)
    hip.synchronize()
    t = time.perf_counter()
    for m in members:
        hip.forward_steps(m["shot"], True, m["V"], coefs, g, src, m["S"], m["upb"], 0, nt)
        hip.backward_steps(m["shot"], True, m["V"], coefs, g, m["S"], m["R"], m["upb"],
                           m["rcv"], m["img"], nt - 1, nt, 1)
    hip.synchronize()
    t_single = time.perf_counter() - t
    log("batch leg: %d shots one at a time: %.2f s" % (nshots, t_single))
    want = [hashlib.sha256(m["img"].download().tobytes()).digest() for m in members]
    reset(members)
    batch = hip.Batch(members, coefs, g)
    hip.synchronize()
    t = time.perf_counter()
    batch.forward_steps(0, nt)
    batch.backward_steps(nt - 1, nt)
    hip.synchronize()
    t_batch = time.perf_counter() - t
    batch.destroy()
    log("batch leg: %d shots batched: %.2f s" % (nshots, t_batch))
    same = all(w == hashlib.sha256(m["img"].download().tobytes()).digest()
               for w, m in zip(want, members))
    # the same batch with FP16 wavefields and upper-border history
    f32img = [m["img"] for m in members]
    for m in members:
        for k in ("S", "R", "upb"):
            m[k].free()
    saved = [hip.get_tiling(c) for c in range(3)]
    for c in range(3):
        hip.set_tiling(c, 0)
    f16 = []
    for m in members:
        shp = (m["shot"].nxe, m["shot"].nye, m["shot"].nze)
        q = dict(m)
        q["shot"] = hip.make_shot(*shp, b, order, nt, nt, (m["shot"].sxe, m["shot"].sye,
                                                         m["shot"].sze), m["shot"].rcv, "f16")
        q["S"], q["R"] = hip.FieldPair(shp, np.uint16), hip.FieldPair(shp, np.uint16)
        q["upb"] = hip.DeviceArray((nt, shp[0], shp[1], order // 2), np.uint16)
        q["img"] = hip.DeviceArray(shp).zero()
        f16.append(q)
    batch = hip.Batch(f16, coefs, g)
    hip.synchronize()
    t = time.perf_counter()
    batch.forward_steps(0, nt)
    batch.backward_steps(nt - 1, nt)
    hip.synchronize()
    t_f16 = time.perf_counter() - t
    batch.destroy()
    log("batch leg: %d shots batched, FP16 wavefields: %.2f s" % (nshots, t_f16))
    num = den = 0.0
    for q, im in zip(f16, f32img):
        a, r = q["img"].download().astype(np.float64), im.download().astype(np.float64)
        num += float(np.sum((a - r) ** 2))
        den += float(np.sum(r ** 2))
    for c, tl in enumerate(saved):
        hip.set_tiling(c, *tl) if tl[0] else hip.set_tiling(c, 0)
    for q, m in zip(f16, members):
        for k in ("S", "R", "upb", "img"):
            q[k].free()
        for k in ("V", "rcv", "img"):
            m[k].free()
    g.free()
    src.free()
    updates = U * (3 * nt - 2)
    full = 2500

    def per_hour(s):  # shots/hour at the workload's 2500 steps, from the per-step cost
        return round(3600.0 * nshots / (s * (3 * full - 2) / (3 * nt - 2)), 1)

    return {
        "workload": "salt1024_batch%d" % nshots, "model": [1024, 1024, 320], "dx_m": d,
        "border": b, "shots": nshots, "nt_run": nt, "nt_workload": full,
        "shot_selection": "all %d seeded windows (nx = ny in 256..448 step 32, "
                          "nz in {256, 288, 320}), each cut around its source" % nshots,
        "extended_grids_cells": round(U / 1e9, 3),
        "single_s": round(t_single, 3), "batched_s": round(t_batch, 3),
        "single_mcell_updates_per_s": round(updates / t_single / 1e6, 1),
        "batched_mcell_updates_per_s": round(updates / t_batch / 1e6, 1),
        "single_shots_per_hour_at_2500_steps": per_hour(t_single),
        "batched_shots_per_hour_at_2500_steps": per_hour(t_batch),
        "speedup": round(t_single / t_batch, 3),
        "images_bit_identical": bool(same),
        "batched_f16_s": round(t_f16, 3),
        "batched_f16_shots_per_hour_at_2500_steps": per_hour(t_f16),
        "f16_speedup_vs_single_f32": round(t_single / t_f16, 3),
        "f16_image_rel_l2_vs_f32": round(float(np.sqrt(num / den)), 6),
        "batch_form": "per-shot stencil launches with the tuned class shapes, small operators "
                      "of all shots in one launch per step, 8-step graph replays",
        "nt_note": "nt reduced to %d of the workload's 2500 to bound the run; the GPU's cost "
                   "per step does not depend on nt, shots/hour are scaled to 2500 steps" % nt,
    }


# ------------------------------------------------------------------- main ----
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nt", type=int, default=3000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--skip", default="",
                    help="comma-separated legs to leave out: cpu, snapshot, "
                         "frequency_domain, three_layer, accuracy, precision, batched")
    ap.add_argument("--batch-nt", type=int, default=300)
    ap.add_argument("--budget", type=float, default=540.0,
                    help="seconds the whole run may take; optional legs that would exceed "
                         "it are skipped")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: the propagation path is measured on a single "
              "MI355X (one device holds the whole model); --gpus %d is not supported"
              % args.gpus, file=sys.stderr)
        sys.exit(2)
    skip = [s for s in args.skip.split(",") if s]
    if args.no_cpu:
        skip.append("cpu")
    budget = Budget(args.budget, skip)

    from rtm3d_demo_amd import hip, host, workloads
    hip.init(0)
    sh = Shot(hip, host, workloads, args.nt)
    log("modeling pass: %.3f s (%.0f Mcell-updates/s)"
        % (sh.modeling_s, sh.U * args.nt / sh.modeling_s / 1e6))
    tiling = None
    if not args.no_tune:
        from rtm3d_demo_amd import autotune
        tshot = hip.make_shot(*sh.shape, sh.b, sh.order, 8, 8, sh.src_ext, sh.geom)
        t = time.time()
        best = autotune.tune(hip, tshot, sh.V, sh.coefs, sh.g, sh.src, sh.F, sh.R, sh.upb,
                             sh.rcv, sh.img, steps=6)
        tiling = autotune.describe(best)
        log("autotune (%.1f s): %s" % (time.time() - t, tiling))
    for i in range(args.warmup):
        sh.migrate()
        hip.synchronize()
        log("warmup shot %d done" % i)
    hip.timing_enable(True)
    hip.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        sh.migrate()
        hip.synchronize()
        log("timed shot %d done at %.2f s" % (i, time.perf_counter() - t0))
    t1 = time.perf_counter()
    cls_t = {c: hip.timing_read(c) for c in range(4)}
    hip.timing_enable(False)
    img = sh.img.download()
    if not np.isfinite(img).all() or not np.abs(img).max() > 0:
        raise RuntimeError("invalid image")
    del img

    shot_s = (t1 - t0) / args.steps
    updates = sh.U * (3 * args.nt - 2)
    value = updates / shot_s / 1e6
    names = {0: ("plain step", 16.0), 1: ("fused receiver step", 28.0),
             2: ("tapered forward step", 16.0),
             3: ("dual backward step (source + receiver + image)", 36.0)}

    # FP32 default (and mode 4): the forward step stores the upper taper of
    # its output, on two-z 128x1x4 waves unless the class tiling asks for
    # one-z 64x1x6 / 64x1x8 waves
    outu = hip.get_taper_fusion("f32") in (-1, 4)
    if outu:
        names[2] = ("tapered forward step (upper taper on the output)", 16.0)

    def cls_stats(c):
        (name, bpc), (n, ms) = names[c], cls_t[c]
        avg = ms / n * 1e-3 if n else None
        tz, by, ry, _ = hip.get_tiling(c)
        if c == 2 and outu and not (tz == 64 and by == 1 and ry in (6, 8)):
            tz, by, ry = 128, 1, 4
        return {"kernel": name, "tile": "%dx%dx%d" % (tz, by, ry) if tz else "default",
                "launches": n, "avg_ms": round(avg * 1e3, 4) if n else None,
                "achieved_GBs": round(bpc * sh.U / avg / 1e9, 1) if n else None,
                "algorithmic_bytes_per_launch": bpc * sh.U}

    # roofline: the class with the most kernel time in the timed region
    dom = max(cls_t, key=lambda c: cls_t[c][1])
    dstat = cls_stats(dom)
    achieved = dstat["achieved_GBs"]
    traffic, tsrc = read_traffic(dom, dstat["tile"])
    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "Mcell-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(shot_s * 1e3, 2), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "hbc_rtm_salt512_nt%d" % args.nt, "grid": [512, 512, 256],
                   "extended_grid": list(sh.shape), "dx_m": 12.5, "border": 40,
                   "nt": args.nt, "stencil_order": 8, "receivers": [128, 128],
                   "shots_per_step": 1, "parallelism": "single_gpu"},
        "shots_per_hour": round(3600.0 / shot_s, 1),
        "modeling_mcell_updates_per_s": round(sh.U * args.nt / sh.modeling_s / 1e6, 1),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "kernel": dstat["kernel"], "kernel_class": dom,
                     "tile": dstat["tile"], "launches": dstat["launches"],
                     "avg_ms": dstat["avg_ms"],
                     "algorithmic_bytes_per_launch": dstat["algorithmic_bytes_per_launch"],
                     "traffic_source": ("profiles/%s (rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                                        "same kernel and tile)" % tsrc) if tsrc else
                     "no committed PMC profile for this class and tile"},
        "stencil_classes": {str(c): cls_stats(c) for c in cls_t if cls_t[c][0]},
        "tiling": tiling,
    }

    if budget.allows("cpu", 45.0):
        try:
            phases = cpu_salt_phases(hip, sh)
            cpu_rate = phases["mcell_updates_per_s"]
            line["cpu_baseline"] = dict(
                {"value": cpu_rate, "unit": "Mcell-updates/s", "cores": cpu_threads(),
                 "kind": "port",
                 "sample": "FP32 oracle port of the reference CPU operators (OpenMP), timed on "
                           "2 steps at each of 9 times of the forward and of the backward pass "
                           "of the salt shot (the GPU's bit-identical states), integrated over "
                           "the 3000-step shot with linear interpolation",
                 "shots_per_hour_extrapolated": round(3600.0 / phases["shot_s_extrapolated"], 3),
                 "gpu_over_cpu": round(value / cpu_rate, 1)}, **host_info(), phases=phases)
        except Exception as e:  # reported, never hidden
            line["cpu_baseline"] = {"error": "%s: %s" % (type(e).__name__, e), "kind": "port",
                                    "cores": cpu_threads()}
    snapshot_ok = budget.allows("snapshot", 20.0)
    if snapshot_ok:
        line["snapshot"] = snapshot_leg(hip, sh, shot_s)
    if budget.allows("frequency_domain", 10.0):
        line["frequency_domain"] = freq_leg(hip, host, sh)
    sh.free()
    gpu_rec = None
    if budget.allows("three_layer", 5.0):
        gpu_rec, line["three_layer_128"] = three_layer_gpu(hip, host, workloads)
    if "cpu_baseline" in line and budget.allows("cpu_three_layer", 100.0):
        c0 = cpu_three_layer_reference(workloads, cpu_threads(),
                                       max(30.0, budget.total - elapsed() - 20.0), gpu_rec)
        line["cpu_baseline"]["three_layer_128_reference"] = c0
        if "mcell_updates_per_s" in c0 and "three_layer_128" in line:
            line["three_layer_128"]["gpu_over_reference_cpu"] = round(
                line["three_layer_128"]["mcell_updates_per_s"] / c0["mcell_updates_per_s"], 1)
    if budget.allows("accuracy", 60.0):
        try:
            acc, snap_acc, case = accuracy_crop(hip, host, workloads)
            line["accuracy_vs_fp64"] = dict(acc, case=case)
            line.setdefault("snapshot", {})["accuracy"] = dict(snap_acc, case=case)
        except Exception as e:
            line["accuracy_vs_fp64"] = {"error": "%s: %s" % (type(e).__name__, e)}
    if budget.allows("precision", 90.0):
        line["precision"] = precision_leg(hip, host, workloads, budget)
    if budget.allows("batched", 90.0):
        line["batched"] = batch_leg(hip, host, workloads, args.batch_nt)
    if budget.skipped:
        line["skipped_legs"] = budget.skipped
    line["run_s"] = round(elapsed(), 1)
    hip.finalize()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
