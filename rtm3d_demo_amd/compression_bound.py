"""What transform coding reaches on the snapshots the learned codec compresses.

The codec stores a 32^3 brick of a source-wavefield snapshot in 1/16 of its
FP16 size.  This script asks how close any coder with that budget can come,
using the best k-term approximation in the orthonormal 3D DCT-II of each brick
(keep the k largest of the 32768 coefficients, exact values, positions for
free -- more than a 1/16 budget really buys), and the same with the budget
pooled over the whole snapshot (the k largest coefficients of all bricks).
For budgets 1/16, 1/8 and 1/4 it rebuilds every snapshot of a snapshot-RTM
run from its kept coefficients, images with them through the library's
snapshot loop, and reports the relative L2 of the image and of the snapshots
against exact FP32 snapshots at the same interval.

Then, with the autoencoder trained as in the benchmark (snapshots of two other
shots of the crop), tiered storage at the 1/16 budget of FP16 snapshots:
per snapshot the bricks ranked by energy, the top E kept in FP16 (16 budget
units each), the next A coded (one unit each: the autoencoder, or 2x2x4 box
decimation in FP16 for comparison), the rest dropped, with 16 E + A equal to
the brick count.

Workload: the 128^3 crop of the 512x512x256 salt model at 12.5 m (border 16,
4000 steps, 25 Hz Ricker at the centre two cells deep, 32x32 receivers every
4 cells), the crop the benchmark's accuracy leg uses; snapshots every 4th
step.  Prints one JSON line.

    python3 rtm3d_demo_amd/compression_bound.py [--nt 4000] [--k 4]
"""
import argparse
import concurrent.futures as cf
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

B = 32


def dct_matrix(n=B):
    """Orthonormal DCT-II: D[u, i] = s(u) cos(pi (2i + 1) u / 2n)."""
    i = np.arange(n)
    d = np.cos(np.pi * (2 * i[None, :] + 1) * i[:, None] / (2 * n)) * np.sqrt(2.0 / n)
    d[0] /= np.sqrt(2.0)
    return d.astype(np.float32)


def to_bricks(region):
    nx, ny, nz = region.shape
    assert nx % B == 0 and ny % B == 0 and nz % B == 0
    r = region.reshape(nx // B, B, ny // B, B, nz // B, B)
    return np.ascontiguousarray(r.transpose(0, 2, 4, 1, 3, 5)).reshape(-1, B, B, B)


def from_bricks(bricks, shape):
    nx, ny, nz = shape
    r = bricks.reshape(nx // B, ny // B, nz // B, B, B, B).transpose(0, 3, 1, 4, 2, 5)
    return np.ascontiguousarray(r).reshape(shape)


def transform(x, d):
    """Separable 3D transform of every brick: y = d applied along each axis."""
    y = x @ d.T                                   # z
    y = (y.transpose(0, 1, 3, 2) @ d.T).transpose(0, 1, 3, 2)   # y
    y = (y.transpose(0, 3, 2, 1) @ d.T).transpose(0, 3, 2, 1)   # x
    return y


def keep_largest(c, keep, pooled):
    """Zero all but the `keep` largest-magnitude coefficients per brick (or
    keep * nbricks over all bricks when pooled)."""
    flat = c.reshape(c.shape[0], -1)
    if pooled:
        a = np.abs(flat).ravel()
        n = keep * flat.shape[0]
        thr = np.partition(a, a.size - n)[a.size - n]
        return np.where(np.abs(c) >= thr, c, 0.0).astype(np.float32)
    idx = np.argpartition(np.abs(flat), flat.shape[1] - keep, axis=1)[:, flat.shape[1] - keep:]
    out = np.zeros_like(flat)
    np.put_along_axis(out, idx, np.take_along_axis(flat, idx, axis=1), axis=1)
    return out.reshape(c.shape)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def ae_train_shots(hip, mk_shot, Vh, coefs, dg, dsrc, nt, shape, b, kt=50):
    """The benchmark's codec: trained on the bricks of the snapshots (every
    50th step after nt/3) of the shots at (32, 96) and (96, 32)."""
    from rtm3d_demo_amd import ae
    lz = shape[2] - 2 * b
    nb = ae.region_bricks((0, 0, 0, shape[0], shape[1], lz))
    slots = [j for j in range((nt - 1) // kt + 1) if nt - 1 - j * kt >= nt // 3]
    pool = hip.DeviceArray((2 * len(slots) * nb, ae.BRICK), np.uint16)
    scl = hip.DeviceArray((nb,))
    row = 0
    for sx, sy in ((32, 96), (96, 32)):
        tshot = mk_shot(sx, sy)
        S = hip.FieldPair(shape)
        snaps = hip.snap_buffer(tshot, "f32", kt)
        hip.snap_forward_steps(tshot, Vh, coefs, dg, dsrc, S, snaps, "f32", kt, 0, nt)
        row = ae.collect_bricks(snaps, slots, pool, row, scl)
        S.free()
        snaps.free()
    codec = ae.Codec(max_bricks=256, seed=3)
    codec.train(pool, row, iters=2000, batch=64, lr=5e-3, seed=5)
    pool.free()
    scl.free()
    return codec


def tiered(hip, codec, exact, approx, image, ref_img, shape, lz, k, nt):
    from rtm3d_demo_amd import ae
    region = (0, 0, 0, shape[0], shape[1], lz)
    nb = ae.region_bricks(region)
    field = hip.DeviceArray((shape[0], shape[1], lz))
    dec = hip.DeviceArray((shape[0], shape[1], lz))
    bricks = hip.DeviceArray((nb, ae.BRICK), np.uint16)
    scl = hip.DeviceArray((nb,))
    codes = hip.DeviceArray((nb, ae.CODES), np.uint16)
    e6 = nb // 16
    e4 = max(0, (nb - 36) // 16)
    variants = [("top %d FP16, rest dropped" % e6, e6, 0, None),
                ("top %d FP16, next %d autoencoder, rest dropped" % (e4, nb - 16 * e4), e4,
                 nb - 16 * e4, "ae"),
                ("top %d FP16, next %d decimated, rest dropped" % (e4, nb - 16 * e4), e4,
                 nb - 16 * e4, "decimate"),
                ("all %d autoencoder" % nb, 0, nb, "ae")]
    out = []
    for name, E, A, kind in variants:
        t = time.perf_counter()
        for j in range(exact.shape[0]):
            snap = exact.slot(j).download()
            eb = to_bricks(snap)
            if kind == "ae":
                field.upload(snap)
                ae.cut_bricks(field, region, bricks, scl)
                codec.encode(bricks, nb, codes)
                codec.decode(codes, scl, nb, dec, region)
                alt = to_bricks(dec.download())
            elif kind == "decimate":
                alt = to_bricks(ae._decimate(snap))
            rank = np.argsort(-(eb.astype(np.float64) ** 2).sum(axis=(1, 2, 3)))
            ob = np.zeros_like(eb)
            ob[rank[:E]] = eb[rank[:E]].astype(np.float16).astype(np.float32)
            if A:
                ob[rank[E:E + A]] = alt[rank[E:E + A]]
            approx.slot(j).upload(from_bricks(ob, snap.shape))
        r = {"budget": "1/16 of FP16 snapshots", "storage": name,
             "image_rel_l2": rel(image(approx), ref_img),
             "seconds": round(time.perf_counter() - t, 1)}
        out.append(r)
        print("[tiered] %s" % json.dumps(r), file=sys.stderr, flush=True)
    for x in (field, dec, bricks, scl, codes):
        x.free()
    codec.destroy()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--nt", type=int, default=4000)
    ap.add_argument("--k", type=int, default=4)
    args = ap.parse_args()
    from rtm3d_demo_amd import hip, host, workloads
    hip.init(0)
    nt, k = args.nt, args.k
    cx = cy = cz = 128
    full = workloads.salt_model()
    vel = np.ascontiguousarray(full[256 - cx // 2:256 + cx // 2, 256 - cy // 2:256 + cy // 2, :cz])
    del full
    d, b, order = 12.5, 16, 8
    nr = (cx // 4, cy // 4)
    p = workloads.params("crop", cx, cy, cz, d, b, nt, workloads.cfl_dt(4500.0, d), 25.0,
                         (cx // 2, cy // 2, 2), (0, 0, 0, 4, 4) + nr)
    geom = hip.RcvGeom(0, 0, 0, 4, 4, *nr)
    coefs = host.stencil_coefs(order, d, d, d)
    V = hip.DeviceArray.from_host(host.v2dt2(host.extend_velocity(vel, b, "abc"), p["dt"]))
    Vh = hip.DeviceArray.from_host(host.v2dt2(host.extend_velocity(vel, b, "hbc"), p["dt"]))
    dsrc = hip.DeviceArray.from_host(host.ricker(nt, p["dt"], p["fpeak"]))
    dg = hip.DeviceArray.from_host(host.taper_table(b, p["taper_factor"]))
    shape = (cx + 2 * b, cy + 2 * b, cz + 2 * b)
    lz = shape[2] - 2 * b
    def mk_shot(sx, sy):
        return hip.make_shot(*shape, b, order, nt, nt, (sx + b, sy + b, 2 + b), geom)

    shot = mk_shot(cx // 2, cy // 2)
    F = hip.FieldPair(shape)
    rcv = hip.DeviceArray(nr + (nt,)).zero()
    hip.model_steps(shot, V, coefs, dg, dsrc, F, rcv, 0, nt)
    F.free()
    exact = hip.snap_buffer(shot, "f32", k)
    S = hip.FieldPair(shape)
    hip.snap_forward_steps(shot, Vh, coefs, dg, dsrc, S, exact, "f32", k, 0, nt)
    S.free()
    nsnap = exact.shape[0]

    def image(snaps):
        R = hip.FieldPair(shape)
        img = hip.DeviceArray(shape).zero()
        hip.snap_backward_steps(shot, Vh, coefs, dg, R, snaps, "f32", k, rcv, img, nt - 1, nt)
        out = img.download()[b:-b, b:-b, b:-b]
        R.free()
        img.free()
        return out

    ref_img = image(exact)
    dm = dct_matrix()
    approx = hip.snap_buffer(shot, "f32", k)
    out = {"case": "128x128x128 crop of the salt model, border 16, %d steps, 25 Hz, "
                   "32x32 receivers; snapshots every %d steps (%d), imaging region "
                   "%dx%dx%d (%d bricks of 32^3)" % (nt, k, nsnap, shape[0], shape[1], lz,
                                                     (shape[0] // B) * (shape[1] // B) * (lz // B)),
           "yardstick": "image from exact FP32 snapshots at the same interval",
           "results": []}
    for ratio in (16, 8, 4):
        for pooled in (False, True):
            t = time.perf_counter()
            keep = B ** 3 // ratio
            errs = []

            def code(snap):
                c = transform(to_bricks(snap), dm)
                return from_bricks(transform(keep_largest(c, keep, pooled), dm.T), snap.shape)

            with cf.ThreadPoolExecutor(8) as pool:
                for j0 in range(0, nsnap, 8):
                    js = range(j0, min(nsnap, j0 + 8))
                    snaps = [exact.slot(j).download() for j in js]
                    for j, snap, rec in zip(js, snaps, pool.map(code, snaps)):
                        if nt - 1 - j * k >= nt // 2 and np.abs(snap).max() > 0:
                            errs.append(rel(rec, snap))
                        approx.slot(j).upload(rec)
            img = image(approx)
            r = {"budget": "1/%d" % ratio,
                 "keep": "%d largest DCT coefficients %s" % (
                     keep, "per snapshot brick on average (pooled over the snapshot)" if pooled
                     else "of each brick"),
                 "image_rel_l2": rel(img, ref_img),
                 "snapshot_rel_l2_mean_second_half": float(np.mean(errs)),
                 "seconds": round(time.perf_counter() - t, 1)}
            out["results"].append(r)
            print("[bound] %s" % json.dumps(r), file=sys.stderr, flush=True)
    out["tiered"] = tiered(hip, ae_train_shots(hip, mk_shot, Vh, coefs, dg, dsrc, nt, shape, b),
                           exact, approx, image, ref_img, shape, lz, k, nt)
    for x in (V, Vh, dsrc, dg, rcv, exact, approx):
        x.free()
    hip.finalize()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
