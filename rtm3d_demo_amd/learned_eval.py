"""Learned snapshot compression on the full salt workload: train on the
snapshots of 32 seeded shots, evaluate on 16 held-out shot positions.

Workload: the 512x512x256 salt model at 12.5 m (HBC border 40, extended
592x592x336), 25 Hz Ricker two cells deep, 128x128 receivers every 4 cells,
nt steps (3000), snapshots every 4th step over the imaging region
(592x592x256: 2888 bricks of 32^3).

Training: for each training shot the forward snapshot pass keeps FP32
snapshots every 150th step; from each snapshot after nt/3, 64 of its non-zero
bricks (seeded draw) go to the training pool; the codec (2048 FP16 codes per
brick) then takes `iters` Adam steps on batches of 64 pool bricks.

Evaluation, per held-out shot: the record is modelled (ABC velocity), then
migrated by snapshot RTM twice through the library's loops -- with FP16
snapshots (rtmhip_snap_*; their image differs from the exact-snapshot image by
about 1e-4, so they serve as the yardstick: FP32 snapshots every 4th step of
this grid would take 269 GB) and with coded snapshots (rtmhip_snap_ae_*) --
and the relative L2 of the coded image against the yardstick image over the
512x512x256 interior is reported, with the times of both migrations.  The
first two held-out shots are also migrated with the untrained network (the
seeded initial weights) to show what training gains.

--codec basis evaluates the transform coder instead (rtmhip_snap_dct_*): its
learned separable basis is trained on FP32 snapshots of the first
--basis-train training shots (every 150th step after nt/3), and each held-out
shot is migrated with the DCT-II and with the learned basis at each
--keep-div budget, against the same FP16-snapshot yardstick.

    python3 rtm3d_demo_amd/learned_eval.py [--nt 3000] [--train 32] [--test 16]
                                           [--iters 4000] [--codec ae|basis]
                                           [--basis-train 8] [--keep-div 32,64]
Prints progress lines to stderr and one JSON line to stdout.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

T0 = time.perf_counter()


def log(msg):
    print("[learned %6.1fs] %s" % (time.perf_counter() - T0, msg), file=sys.stderr, flush=True)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def positions(seed, n, taken=()):
    """n distinct seeded source positions (x, y) on a 16-cell lattice inside
    [32, 480]^2, none in `taken`."""
    rng = np.random.default_rng(seed)
    cand = [(x, y) for x in range(32, 481, 16) for y in range(32, 481, 16) if (x, y) not in taken]
    idx = rng.choice(len(cand), n, replace=False)
    return [cand[i] for i in idx]


def basis_eval(c, keep_divs):
    """Transform-coded snapshot RTM of the held-out shots with the DCT-II and
    with a basis learned from the training shots' snapshots."""
    from rtm3d_demo_amd.basis import BasisLearner
    hip, shape, b, nt, k = c["hip"], c["shape"], c["b"], c["nt"], c["k"]
    V, coefs, g, src = c["V"], c["coefs"], c["g"], c["src"]
    kt = 150
    slots = [j for j in range((nt - 1) // kt + 1) if nt - 1 - j * kt >= nt // 3]
    learner = BasisLearner()
    F, R = hip.FieldPair(shape), hip.FieldPair(shape)
    t = time.perf_counter()
    for i, pos in enumerate(c["train"]):
        shot = c["mk"](pos)
        for f in (F.a, F.b):
            f.zero()
        F.p.value, F.pp.value = F.a.ptr.value, F.b.ptr.value
        snaps = hip.snap_buffer(shot, "f32", kt)
        hip.snap_forward_steps(shot, V, coefs, g, src, F, snaps, "f32", kt, 0, nt)
        for j in slots:
            learner.add(snaps.slot(j).download(), b)
        snaps.free()
        log("basis training shot %d/%d at %s" % (i + 1, len(c["train"]), pos))
    basis = learner.basis()
    train_s = time.perf_counter() - t
    img = hip.DeviceArray(shape).zero()
    rcv = hip.DeviceArray(c["nr"] + (nt,))
    fp16 = hip.snap_buffer(c["mk"](c["test"][0]), "f16", k)

    def migrate(shot, div=None, bas=None):
        for f in (F.a, F.b, R.a, R.b, img):
            f.zero()
        F.p.value, F.pp.value = F.a.ptr.value, F.b.ptr.value
        R.p.value, R.pp.value = R.a.ptr.value, R.b.ptr.value
        hip.synchronize()
        t = time.perf_counter()
        if div is None:
            hip.snap_forward_steps(shot, V, coefs, g, src, F, fp16, "f16", k, 0, nt)
            hip.snap_backward_steps(shot, V, coefs, g, R, fp16, "f16", k, rcv, img, nt - 1, nt)
        else:
            st = hip.DctSnapshots(shot, k, div, bas)
            st.forward_steps(V, coefs, g, src, F, 0, nt)
            st.backward_steps(V, coefs, g, R, rcv, img, nt - 1, nt)
            st.free()
        hip.synchronize()
        s = time.perf_counter() - t
        return img.download()[b:-b, b:-b, b:-b], s

    results = []
    for i, pos in enumerate(c["test"]):
        shot = c["mk"](pos)
        for f in (F.a, F.b):
            f.zero()
        F.p.value, F.pp.value = F.a.ptr.value, F.b.ptr.value
        rcv.zero()
        hip.model_steps(shot, c["V_abc"], coefs, g, src, F, rcv, 0, nt)
        ref, s_ref = migrate(shot)
        r = {"source": list(pos), "fp16_snapshot_rtm_s": round(s_ref, 3)}
        for div in keep_divs:
            for name, bas in (("dct", None), ("learned", basis)):
                got, s = migrate(shot, div, bas)
                r["%s_keep%d_image_rel_l2" % (name, div)] = rel(got, ref)
                r["%s_keep%d_rtm_s" % (name, div)] = round(s, 3)
        results.append(r)
        log("held-out shot %d/%d: %s" % (i + 1, len(c["test"]), json.dumps(r)))
    st = hip.DctSnapshots(c["mk"](c["test"][0]), k, keep_divs[0])
    summary = {}
    for key in results[0]:
        if key.endswith("image_rel_l2"):
            v = [r[key] for r in results]
            summary[key] = {"mean": float(np.mean(v)), "max": float(np.max(v))}
    out = {
        "workload": "salt 512x512x256 at 12.5 m, HBC border 40, %d steps, snapshots every %d "
                    "steps, transform-coded 8x8x32 bricks" % (nt, k),
        "training": {"shots": len(c["train"]), "snapshots": len(c["train"]) * len(slots),
                     "seconds": round(train_s, 1)},
        "held_out": results, "summary": summary,
        "coded_GB_per_shot_keep%d" % keep_divs[0]: round(st.nbytes() / 1e9, 2),
        "fp16_snapshot_GB_per_shot": round(fp16.nbytes / 1e9, 2),
        "yardstick": "snapshot RTM with FP16 snapshots at the same interval",
    }
    st.free()
    for x in (F, R, img, rcv, fp16):
        x.free()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--nt", type=int, default=3000)
    ap.add_argument("--train", type=int, default=32)
    ap.add_argument("--test", type=int, default=16)
    ap.add_argument("--iters", type=int, default=4000)
    ap.add_argument("--k", type=int, default=4)
    ap.add_argument("--codec", choices=("ae", "basis"), default="ae")
    ap.add_argument("--basis-train", type=int, default=8)
    ap.add_argument("--keep-div", default="32,64")
    args = ap.parse_args()
    from rtm3d_demo_amd import ae, hip, host, workloads
    hip.init(0)
    nt, k = args.nt, args.k
    p = workloads.salt_512_params(nt)
    b, order = p["blen"], p["stencil_order"]
    vel = workloads.salt_model()
    coefs = host.stencil_coefs(order, p["dx"], p["dy"], p["dz"])
    V_abc = hip.DeviceArray.from_host(host.v2dt2(host.extend_velocity(vel, b, "abc"), p["dt"]))
    V = hip.DeviceArray.from_host(host.v2dt2(host.extend_velocity(vel, b, "hbc"), p["dt"]))
    shape = tuple(n + 2 * b for n in vel.shape)
    del vel
    src = hip.DeviceArray.from_host(host.ricker(nt, p["dt"], p["fpeak"]))
    g = hip.DeviceArray.from_host(host.taper_table(b, p["taper_factor"]))
    geom = hip.RcvGeom(p["receiver_start_x"], p["receiver_start_y"], p["receiver_depth_z"],
                       p["receiver_distance_x"], p["receiver_distance_y"],
                       p["receiver_count_x"], p["receiver_count_y"])
    nr = (p["receiver_count_x"], p["receiver_count_y"])

    def mk(pos):
        return hip.make_shot(*shape, b, order, nt, nt, (pos[0] + b, pos[1] + b, 2 + b), geom)

    train = positions(11, args.train)
    test = positions(23, args.test, set(train))
    lz = shape[2] - 2 * b
    region = (0, 0, 0, shape[0], shape[1], lz)
    nb = ae.region_bricks(region)
    L = hip.load()
    if args.codec == "basis":
        ctx = dict(hip=hip, V=V, V_abc=V_abc, coefs=coefs, g=g, src=src, shape=shape, b=b,
                   nt=nt, k=k, nr=nr, mk=mk, train=train[:args.basis_train], test=test)
        out = basis_eval(ctx, [int(d) for d in args.keep_div.split(",")])
        for x in (V_abc, V, src, g):
            x.free()
        hip.finalize()
        print(json.dumps(out), flush=True)
        return 0

    # ---- training pool
    kt, per = 150, 64
    slots = [j for j in range((nt - 1) // kt + 1) if nt - 1 - j * kt >= nt // 3]
    pool = hip.DeviceArray((len(train) * len(slots) * per, ae.BRICK), np.uint16)
    tmp = hip.DeviceArray((nb, ae.BRICK), np.uint16)
    scl = hip.DeviceArray((nb,))
    rng = np.random.default_rng(5)
    row = 0
    F = hip.FieldPair(shape)
    for i, pos in enumerate(train):
        shot = mk(pos)
        for f in (F.a, F.b):
            f.zero()
        F.p.value, F.pp.value = F.a.ptr.value, F.b.ptr.value
        snaps = hip.snap_buffer(shot, "f32", kt)
        hip.snap_forward_steps(shot, V, coefs, g, src, F, snaps, "f32", kt, 0, nt)
        for j in slots:
            ae.cut_bricks(snaps.slot(j), region, tmp, scl)
            nz = np.flatnonzero(scl.download() > 0)
            for bi in rng.choice(nz, min(per, nz.size), replace=False):
                hip.check(L.rtmhip_copy_device(pool.view(row, 1).ptr, tmp.view(int(bi), 1).ptr,
                                               ae.BRICK * 2))
                row += 1
        snaps.free()
        log("training shot %d/%d at %s: pool %d bricks" % (i + 1, len(train), pos, row))
    tmp.free()
    scl.free()
    codec = ae.Codec(max_bricks=1024, seed=3)
    untrained = codec.params()
    t = time.perf_counter()
    losses = codec.train(pool, row, iters=args.iters, batch=64, lr=5e-3, seed=7)
    hip.synchronize()
    train_s = time.perf_counter() - t
    pool.free()
    log("trained: mse %.3e -> %.3e in %.1f s" % (losses[:20].mean(), losses[-50:].mean(), train_s))
    codec0 = ae.Codec(max_bricks=1024, seed=3)
    codec0.set_params(untrained)

    # ---- held-out shots
    R = hip.FieldPair(shape)
    img = hip.DeviceArray(shape).zero()
    rcv = hip.DeviceArray(nr + (nt,))
    fp16 = hip.snap_buffer(mk(test[0]), "f16", k)

    def migrate(shot, coded_with=None):
        for f in (F.a, F.b, R.a, R.b, img):
            f.zero()
        F.p.value, F.pp.value = F.a.ptr.value, F.b.ptr.value
        R.p.value, R.pp.value = R.a.ptr.value, R.b.ptr.value
        hip.synchronize()
        t = time.perf_counter()
        if coded_with is None:
            hip.snap_forward_steps(shot, V, coefs, g, src, F, fp16, "f16", k, 0, nt)
            hip.snap_backward_steps(shot, V, coefs, g, R, fp16, "f16", k, rcv, img, nt - 1, nt)
        else:
            cs = ae.CodedSnapshots(shot, k, coded_with)
            cs.forward_steps(V, coefs, g, src, F, 0, nt)
            cs.backward_steps(V, coefs, g, R, rcv, img, nt - 1, nt)
            cs.free()
        hip.synchronize()
        s = time.perf_counter() - t
        return img.download()[b:-b, b:-b, b:-b], s

    results = []
    for i, pos in enumerate(test):
        shot = mk(pos)
        for f in (F.a, F.b):
            f.zero()
        F.p.value, F.pp.value = F.a.ptr.value, F.b.ptr.value
        rcv.zero()
        hip.model_steps(shot, V_abc, coefs, g, src, F, rcv, 0, nt)
        ref, s_ref = migrate(shot)
        got, s_ae = migrate(shot, codec)
        r = {"source": list(pos), "image_rel_l2": rel(got, ref), "fp16_snapshot_rtm_s": round(s_ref, 3),
             "coded_snapshot_rtm_s": round(s_ae, 3)}
        if i < 2:
            raw, _ = migrate(shot, codec0)
            r["untrained_image_rel_l2"] = rel(raw, ref)
        results.append(r)
        log("held-out shot %d/%d: %s" % (i + 1, len(test), json.dumps(r)))
    errs = [r["image_rel_l2"] for r in results]
    cs = ae.CodedSnapshots(mk(test[0]), k, codec)
    out = {
        "workload": "salt 512x512x256 at 12.5 m, HBC border 40, %d steps, snapshots every %d "
                    "steps over 592x592x256 (%d bricks of 32^3)" % (nt, k, nb),
        "training": {"shots": len(train), "pool_bricks": row, "iters": args.iters, "batch": 64,
                     "lr": 5e-3, "mse_first_20": float(losses[:20].mean()),
                     "mse_last_50": float(losses[-50:].mean()), "seconds": round(train_s, 1)},
        "held_out": results,
        "image_rel_l2_mean": float(np.mean(errs)), "image_rel_l2_max": float(np.max(errs)),
        "coded_GB_per_shot": round(cs.nbytes() / 1e9, 2),
        "fp16_snapshot_GB_per_shot": round(fp16.nbytes / 1e9, 2),
        "yardstick": "snapshot RTM with FP16 snapshots at the same interval",
    }
    cs.free()
    for x in (V_abc, V, src, g, F, R, img, rcv, fp16):
        x.free()
    codec.destroy()
    codec0.destroy()
    hip.finalize()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
