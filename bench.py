#!/usr/bin/env python3
"""Benchmark of the RTM propagation-and-imaging path on one MI355X.

Workload (one "step" = one complete FP32 HBC RTM shot: forward propagation
with upper-border history, backward source reconstruction, receiver
back-propagation and cross-correlation imaging):
  512x512x256 salt model at 12.5 m (v(z) gradient, seeded smoothed +-5 %
  perturbation, seeded ellipsoidal salt at 4500 m/s), HBC border 40
  (extended 592x592x336), 3000 steps, 25 Hz Ricker, 128x128 surface receivers
  (stride 4); the shot record is synthesised first by the modeling path.

Reported: Mcell-updates/s of the whole shot (updated cells per wavefield step
x (3 nt - 2) wavefield steps / shot time), RTM shots/hour, the roofline of the
dominant kernel (the stencil step: 16 algorithmic bytes per updated cell,
durations from HIP events around every launch in the timed region), and the
reference CPU program timed on a bounded sample of the same workload on this
host's cores.  Beside the FP32 line: the same shot with FP16 and with BF16
wavefield storage (FP32 arithmetic; 10 algorithmic bytes per stencil cell);
every format's trace / image error against the FP64 oracle on a 128^3 crop
over 4000 steps; and a batched leg (8 shots with uneven apertures cut from a
1024x1024x320 salt model, 2500 steps, batched against one at a time, images
compared bit for bit).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
       [--no-cpu] [--no-precision] [--no-tune] [--batch-shots N] [--batch-nt T]
"""
import argparse
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


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


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
        self.src = hip.DeviceArray.from_host(host.ricker(nt, p["dt"], p["fpeak"]))
        self.g = hip.DeviceArray.from_host(host.taper_table(b, p["taper_factor"]))
        self.geom = hip.RcvGeom(p["receiver_start_x"], p["receiver_start_y"],
                                p["receiver_depth_z"], p["receiver_distance_x"],
                                p["receiver_distance_y"], p["receiver_count_x"],
                                p["receiver_count_y"])
        self.shot = hip.make_shot(*self.shape, b, order, nt, nt,
                                  (p["source_start_x"] + b, p["source_start_y"] + b,
                                   p["source_depth_z"] + b), self.geom)
        self.F = hip.FieldPair(self.shape)
        self.R = hip.FieldPair(self.shape)
        self.extra = {}  # 16-bit wavefield buffers per format
        self.rcv = hip.DeviceArray((p["receiver_count_x"], p["receiver_count_y"], nt)).zero()
        self.img = hip.DeviceArray(self.shape).zero()
        self.upb = hip.DeviceArray((nt, self.shape[0], self.shape[1], order // 2))
        # modeling (ABC velocity) synthesises the record; migration uses HBC
        t = time.time()
        hip.model_steps(self.shot, self.V, self.coefs, self.g, self.src, self.F, self.rcv, 0, nt)
        hip.synchronize()
        self.modeling_s = time.time() - t
        self.V.upload(host.v2dt2(host.extend_velocity(vel, b, "hbc"), p["dt"]))
        del vel
        self.U = float((self.shape[0] - order) * (self.shape[1] - order) * (self.shape[2] - order))

    def buffers(self, fmt):
        """(shot, F, R, upb) for a wavefield storage format."""
        if fmt == "f32":
            return self.shot, self.F, self.R, self.upb
        if fmt not in self.extra:
            hip, p, b = self.hip, self.p, self.b
            dt = hip.WAVE_FORMATS[fmt][1]
            shot = hip.make_shot(*self.shape, b, self.order, self.nt, self.nt,
                                 (p["source_start_x"] + b, p["source_start_y"] + b,
                                  p["source_depth_z"] + b), self.geom, fmt)
            self.extra[fmt] = (shot, hip.FieldPair(self.shape, dt), hip.FieldPair(self.shape, dt),
                               hip.DeviceArray((self.nt, self.shape[0], self.shape[1],
                                                self.order // 2), dt))
        return self.extra[fmt]

    def release(self, fmt):
        if fmt in self.extra:
            for x in self.extra.pop(fmt)[1:]:
                x.free()

    def migrate(self, fmt="f32"):
        hip = self.hip
        shot, F, R, upb = self.buffers(fmt)
        for f in (F.a, F.b, R.a, R.b, self.img):
            f.zero()
        hip.forward_steps(shot, True, self.V, self.coefs, self.g, self.src, F, upb, 0, self.nt)
        # mode 1: fused receiver step (the one-pass source + receiver mode 2
        # is bit-identical but, with its 4-row tiles, not faster here)
        hip.backward_steps(shot, True, self.V, self.coefs, self.g, F, R, upb, self.rcv, self.img,
                           self.nt - 1, self.nt, 1)


def read_traffic(cls):
    """HBM bytes per launch of a stencil kernel class, from the committed
    rocprofv3 PMC summary ("traffic_bytes_per_launch[class N]: bytes")."""
    path = os.path.join(REPO, "profiles", "stencil_salt512_pmc.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch\[class %d\]:\s*([0-9.eE+]+)" % cls,
                  open(path).read())
    return float(m.group(1)) if m else None


# ---------------------------------------------------------------- CPU leg --
def _ref_report(stdout, key):
    m = re.search(r'"%s":\s*([0-9.eE+-]+)' % key, stdout)
    return float(m.group(1)) if m else None


def cpu_reference_sample(workloads, steps, cores):
    """The reference CPU program (compiled from its sources) on the salt model
    for `steps` time steps: modeling, then HBC migration of that record."""
    wd = tempfile.mkdtemp(prefix="rtmcpu_")
    try:
        for d in ("vel", "out", "tmp", "dat"):
            os.makedirs(os.path.join(wd, d))
        workloads.salt_model().tofile(os.path.join(wd, "vel", "model.bin"))
        p = workloads.salt_512_params(steps)
        p["nthreads"] = cores
        out = {}
        for mode in ("modeling", "migration"):
            q = dict(p, modeling=(mode == "modeling"), migration=(mode == "migration"))
            with open(os.path.join(wd, "in.json"), "w") as f:
                json.dump(q, f)
            t = time.time()
            r = subprocess.run([REFBIN, "in.json"], cwd=wd, capture_output=True, text=True,
                               timeout=900, env=dict(os.environ, OMP_NUM_THREADS=str(cores)))
            wall = time.time() - t
            if r.returncode != 0:
                raise RuntimeError("reference run failed: " + r.stdout[-500:])
            out[mode] = {"wall_s": wall,
                         "kernel_s": _ref_report(r.stdout, "%s_time" % mode)}
            log("reference %s: %s" % (mode, out[mode]))
        return out
    finally:
        shutil.rmtree(wd, ignore_errors=True)


def accuracy_crop(hip, host, workloads, formats=("f32",), crop=(80, 80, 64), nt=400):
    """Device results vs the FP64 oracle on a crop of the salt model
    (default 80x80x64 + border 16, 400 steps): relative L2 of the receiver
    traces and of the migrated HBC image, for each wavefield storage format
    (each format models its own record and migrates it)."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import pyoracle as po

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
    results = {}
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
        out = hip.DeviceArray(crop)
        hip.image_finish(img, out, b, host.stencil_coefs(2, d, d, d))
        results[fmt] = (rec, out.download())
        for x in (F, rcv, S, R, upb, img, out):
            x.free()
    for x in (V, Vh, dsrc, dg):
        x.free()
    rec64 = po.pipeline_modeling(p, vel, np.float64)[0, 0]
    q = dict(p, modeling=False, migration=True)
    img64 = po.pipeline_migration(q, vpe, rec64[None, None], np.float64)
    rel = lambda a, r: float(np.linalg.norm(a - r) / np.linalg.norm(r))
    return {fmt: {"trace_rel_l2": rel(rec.astype(np.float64), rec64),
                  "image_rel_l2": rel(im.astype(np.float64), img64)}
            for fmt, (rec, im) in results.items()}, \
        "%dx%dx%d crop of the salt model, border 16, %d steps, %dx%d receivers" % (
            cx, cy, cz, nt, nr[0], nr[1])


class BatchShots:
    """Device state of a batch of shots of the 1024 x 1024 x 320 model: each
    shot's window (uneven apertures), its record synthesised by the modeling
    path, and the buffers of its migration."""

    def __init__(self, hip, host, workloads, nshots, nt):
        self.hip, self.nt = hip, nt
        wins = workloads.salt_1024_shots(64)[:nshots]
        p0 = workloads.salt_1024_params(wins[0][2], wins[0][3], wins[0][4], nt)
        b, order = p0["blen"], p0["stencil_order"]
        self.b = b
        self.coefs = host.stencil_coefs(order, p0["dx"], p0["dy"], p0["dz"])
        self.g = hip.DeviceArray.from_host(host.taper_table(b, p0["taper_factor"]))
        self.src = hip.DeviceArray.from_host(host.ricker(nt, p0["dt"], p0["fpeak"]))
        self.members, self.grids = [], []
        self.U = 0.0
        t = time.time()
        for x0, y0, nxy, nz, src in wins:
            p = workloads.salt_1024_params(nxy, nz, src, nt)
            vel = workloads.salt_1024_window(x0, y0, nxy, nxy, nz)
            ve = host.extend_velocity(vel, b, "abc")
            shape = ve.shape
            V = hip.DeviceArray.from_host(host.v2dt2(ve, p["dt"]))
            geom = hip.RcvGeom(0, 0, 0, 4, 4, nxy // 4, nxy // 4)
            shot = hip.make_shot(*shape, b, order, nt, nt, (src[0] + b, src[1] + b, src[2] + b),
                                 geom)
            m = dict(shot=shot, V=V, ricker=self.src, S=hip.FieldPair(shape),
                     R=hip.FieldPair(shape),
                     rcv=hip.DeviceArray((nxy // 4, nxy // 4, nt)).zero(),
                     img=hip.DeviceArray(shape).zero(),
                     upb=hip.DeviceArray((nt, shape[0], shape[1], order // 2)))
            hip.model_steps(shot, V, self.coefs, self.g, self.src, m["S"], m["rcv"], 0, nt)
            V.upload(host.v2dt2(host.extend_velocity(vel, b, "hbc"), p["dt"]))
            self.members.append(m)
            self.grids.append(list(shape))
            self.U += float((shape[0] - order) * (shape[1] - order) * (shape[2] - order))
        hip.synchronize()
        log("batch: %d shot windows built and modelled in %.1f s" % (nshots, time.time() - t))

    def reset(self):
        for m in self.members:
            for f in (m["S"], m["R"]):
                f.a.zero()
                f.b.zero()
                f.p.value, f.pp.value = f.a.ptr.value, f.b.ptr.value
            m["img"].zero()

    def single(self):
        """The shots one at a time through the per-shot loops (plain FP32
        single-shot GPU RTM)."""
        hip, nt = self.hip, self.nt
        for m in self.members:
            hip.forward_steps(m["shot"], True, m["V"], self.coefs, self.g, self.src, m["S"],
                              m["upb"], 0, nt)
            hip.backward_steps(m["shot"], True, m["V"], self.coefs, self.g, m["S"], m["R"],
                               m["upb"], m["rcv"], m["img"], nt - 1, nt, 1)

    def free(self):
        for m in self.members:
            for k in ("V", "S", "R", "rcv", "img", "upb"):
                m[k].free()
        self.g.free()
        self.src.free()


def batch_leg(hip, host, workloads, nshots, nt, tune=True):
    """Batched RTM of shots of the 1024 x 1024 x 320 model against the same
    shots run one at a time (plain single-shot GPU RTM, with the per-shot
    tile shapes the batch tuner found best on these grids); images compared
    bit for bit."""
    from rtm3d_demo_amd import autotune

    bs = BatchShots(hip, host, workloads, nshots, nt)
    modes = None
    batch = hip.Batch(bs.members, bs.coefs, bs.g)
    if tune:
        t = time.time()
        modes = autotune.tune_batch(hip, batch, nt, log=log)
        log("batch autotune (%.1f s): %s" % (time.time() - t, modes))
    chosen = {c: modes[c]["mode"] for c in modes} if modes else {}
    batch.destroy()
    bs.reset()
    hip.synchronize()
    t = time.perf_counter()
    bs.single()
    hip.synchronize()
    t_single = time.perf_counter() - t
    log("batch leg: %d shots one at a time: %.2f s" % (nshots, t_single))
    want = [m["img"].download() for m in bs.members]
    bs.reset()
    batch = hip.Batch(bs.members, bs.coefs, bs.g)
    for c, m in chosen.items():
        batch.set_stencil(c, m)
    hip.synchronize()
    t = time.perf_counter()
    batch.forward_steps(0, nt)
    batch.backward_steps(nt - 1, nt)
    hip.synchronize()
    t_batch = time.perf_counter() - t
    log("batch leg: %d shots batched: %.2f s" % (nshots, t_batch))
    same = all(np.array_equal(w.view(np.uint32), m["img"].download().view(np.uint32))
               for w, m in zip(want, bs.members))
    batch.destroy()
    # the same batch with FP16 wavefields and UPB history (FP32 arithmetic,
    # velocity, records and images), default 16-bit shapes, image error
    # against the FP32 images
    saved = [hip.get_tiling(c) for c in range(3)]
    for c in range(3):
        hip.set_tiling(c, 0)
    f16 = []
    for m in bs.members:
        shp = (m["shot"].nxe, m["shot"].nye, m["shot"].nze)
        q = dict(m)
        q["shot"] = hip.make_shot(*shp, bs.b, 8, nt, nt, (m["shot"].sxe, m["shot"].sye,
                                                         m["shot"].sze), m["shot"].rcv, "f16")
        q["S"], q["R"] = hip.FieldPair(shp, np.uint16), hip.FieldPair(shp, np.uint16)
        q["upb"] = hip.DeviceArray(m["upb"].shape, np.uint16)
        q["img"] = m["img"]
        m["img"].zero()
        f16.append(q)
    batch = hip.Batch(f16, bs.coefs, bs.g)
    hip.synchronize()
    t = time.perf_counter()
    batch.forward_steps(0, nt)
    batch.backward_steps(nt - 1, nt)
    hip.synchronize()
    t_f16 = time.perf_counter() - t
    batch.destroy()
    log("batch leg: %d shots batched, FP16 wavefields: %.2f s" % (nshots, t_f16))
    err16 = float(np.sqrt(sum(float(np.sum((q["img"].download().astype(np.float64) - w) ** 2))
                              for q, w in zip(f16, want)) /
                          sum(float(np.sum(w.astype(np.float64) ** 2)) for w in want)))
    for q in f16:
        for k in ("S", "R", "upb"):
            q[k].free()
    for c, tl in enumerate(saved):
        hip.set_tiling(c, *tl) if tl[0] else hip.set_tiling(c, 0)
    del want
    updates = bs.U * (3 * nt - 2)
    out = {
        "workload": "salt1024_batch%d_nt%d" % (nshots, nt),
        "model": [1024, 1024, 320], "dx_m": 12.5, "border": 40, "nt": nt, "shots": nshots,
        "shot_selection": "the first %d of 64 seeded windows (nx = ny in 256..448 step 32, "
                          "nz in {256, 288, 320})" % nshots,
        "extended_grids": bs.grids,
        "batched_s": round(t_batch, 3), "single_s": round(t_single, 3),
        "batched_mcell_updates_per_s": round(updates / t_batch / 1e6, 1),
        "single_mcell_updates_per_s": round(updates / t_single / 1e6, 1),
        "batched_shots_per_hour": round(3600.0 * nshots / t_batch, 1),
        "single_shots_per_hour": round(3600.0 * nshots / t_single, 1),
        "speedup": round(t_single / t_batch, 3),
        "images_bit_identical": bool(same),
        "batched_f16_s": round(t_f16, 3),
        "batched_f16_shots_per_hour": round(3600.0 * nshots / t_f16, 1),
        "f16_speedup_vs_single_f32": round(t_single / t_f16, 3),
        "f16_image_rel_l2_vs_f32": round(err16, 6),
        "autotune": {str(c): v for c, v in sorted(modes.items())} if modes else None,
        "graph_steps": 8,
    }
    bs.free()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nt", type=int, default=3000)
    ap.add_argument("--cpu-steps", type=int, default=6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the FP16 / BF16 wavefield-storage shots")
    ap.add_argument("--batch-shots", type=int, default=8,
                    help="shots of the batched leg (0 skips it)")
    ap.add_argument("--batch-nt", type=int, default=2500)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: the propagation path is measured on a single "
              "MI355X (one device holds the whole model); --gpus %d is not supported"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    from rtm3d_demo_amd import hip, host, workloads
    hip.init(0)
    sh = Shot(hip, host, workloads, args.nt)
    log("modeling pass: %.3f s (%.0f Mcell-updates/s)"
        % (sh.modeling_s, sh.U * args.nt / sh.modeling_s / 1e6))
    tiling = None
    if not args.no_tune:
        from rtm3d_demo_amd import autotune
        p = sh.p
        tshot = hip.make_shot(*sh.shape, sh.b, sh.order, 8, 8,
                              (p["source_start_x"] + sh.b, p["source_start_y"] + sh.b,
                               p["source_depth_z"] + sh.b), sh.geom)
        t = time.time()
        best = autotune.tune(hip, tshot, sh.V, sh.coefs, sh.g, sh.src, sh.F, sh.R, sh.upb,
                             sh.rcv, sh.img, steps=6, log=log)
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
    n0, ms0 = hip.timing_read(0)
    n1, ms1 = hip.timing_read(1)
    n2, ms2 = hip.timing_read(2)
    n3, ms3 = hip.timing_read(3)
    hip.timing_enable(False)
    img = sh.img.download()
    if not np.isfinite(img).all() or not np.abs(img).max() > 0:
        raise RuntimeError("invalid image")
    del img

    shot_s = (t1 - t0) / args.steps
    updates = sh.U * (3 * args.nt - 2)
    value = updates / shot_s / 1e6
    # stencil kernel classes in the timed region: launches, total ms and
    # algorithmic bytes per updated cell
    classes = {0: ("plain step", n0, ms0, 16.0), 1: ("fused receiver step", n1, ms1, 28.0),
               2: ("tapered forward step", n2, ms2, 16.0),
               3: ("dual backward step (source + receiver + image)", n3, ms3, 36.0)}

    def cls_stats(c):
        name, n, ms, bpc = classes[c]
        avg = ms / n * 1e-3 if n else None
        return {"kernel": name, "launches": n, "avg_ms": round(avg * 1e3, 4) if n else None,
                "achieved_GBs": round(bpc * sh.U / avg / 1e9, 1) if n else None,
                "algorithmic_bytes_per_launch": bpc * sh.U}

    # roofline: the class with the most kernel time in the timed region
    dom = max(classes, key=lambda c: classes[c][2])
    dstat = cls_stats(dom)
    achieved = dstat["achieved_GBs"]
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
                     "traffic": read_traffic(dom), "kernel": dstat["kernel"],
                     "kernel_class": dom, "launches": dstat["launches"],
                     "avg_ms": dstat["avg_ms"],
                     "algorithmic_bytes_per_launch": dstat["algorithmic_bytes_per_launch"]},
        "stencil_classes": {str(c): cls_stats(c) for c in classes if classes[c][1]},
        "tiling": tiling,
    }
    # 16-bit wavefield storage: the same shot, one warm-up and one timed shot
    # per format (arithmetic stays FP32; 10 algorithmic bytes per stencil cell)
    formats = ["f32"]
    if not args.no_precision:
        prec = {}
        # 16-bit storage runs its own default shapes (the FP32 autotune's
        # choices are for 4-byte lanes)
        saved = [hip.get_tiling(c) for c in range(3)]
        for c in range(3):
            hip.set_tiling(c, 0)
        for fmt in ("f16", "bf16"):
            sh.migrate(fmt)
            hip.synchronize()
            hip.timing_enable(True)
            t = time.perf_counter()
            sh.migrate(fmt)
            hip.synchronize()
            s16 = time.perf_counter() - t
            k0, m0 = hip.timing_read(0)
            hip.timing_read(1)
            hip.timing_read(2)
            hip.timing_read(3)
            hip.timing_enable(False)
            sh.release(fmt)
            a0 = m0 / max(k0, 1) * 1e-3
            prec[fmt] = {"ms_per_shot": round(s16 * 1e3, 2),
                         "value": round(updates / s16 / 1e6, 1),
                         "shots_per_hour": round(3600.0 / s16, 1),
                         "speedup_vs_f32": round(shot_s / s16, 3),
                         "stencil_avg_ms": round(a0 * 1e3, 4),
                         "stencil_achieved_GBs": round(10.0 * sh.U / a0 / 1e9, 1),
                         "algorithmic_bytes_per_launch": 10.0 * sh.U}
            log("%s shot: %.3f s (%.2fx FP32)" % (fmt, s16, shot_s / s16))
            formats.append(fmt)
        for c, t in enumerate(saved):
            hip.set_tiling(c, *t) if t[0] else hip.set_tiling(c, 0)
        line["wave_formats"] = prec
    for x in (sh.V, sh.src, sh.g, sh.F, sh.R, sh.rcv, sh.img, sh.upb):
        x.free()
    if args.batch_shots > 0:
        line["batched"] = batch_leg(hip, host, workloads, args.batch_shots, args.batch_nt,
                                    tune=not args.no_tune)

    if not args.no_cpu:
        cores = min(16, os.cpu_count() or 1)
        cpu = None
        try:
            # the mixed-precision configuration's accuracy case: a 128^3 crop,
            # 4000 steps, against the FP64 oracle on the host cores
            acc, case = accuracy_crop(hip, host, workloads, tuple(formats), (128, 128, 128), 4000)
            line["accuracy_vs_fp64"] = dict(acc["f32"], case=case)
            for fmt in formats[1:]:
                line["wave_formats"][fmt].update(acc[fmt])
        except Exception as e:  # reported, never hidden
            line["accuracy_vs_fp64"] = {"error": str(e)}
        if os.path.exists(REFBIN):
            try:
                s = args.cpu_steps
                r = cpu_reference_sample(workloads, s, cores)
                mig = r["migration"]["kernel_s"] or r["migration"]["wall_s"]
                cpu_rate = sh.U * (3 * s - 2) / mig / 1e6
                cpu = {"value": round(cpu_rate, 2), "unit": "Mcell-updates/s", "cores": cores,
                       "kind": "reference",
                       "sample": "reference CPU program, salt model 512x512x256 + HBC 40, "
                                 "%d-step HBC migration (forward+backward+imaging) after a "
                                 "%d-step modeling run; early steps (small wavefield)" % (s, s),
                       "migration_s": round(mig, 3),
                       "shots_per_hour_extrapolated": round(
                           3600.0 / (mig * (3 * args.nt - 2) / (3 * s - 2)), 3),
                       "gpu_over_cpu": round(value / cpu_rate, 1)}
            except Exception as e:
                cpu = {"error": str(e), "kind": "reference", "cores": cores}
        line["cpu_baseline"] = cpu
    hip.finalize()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
