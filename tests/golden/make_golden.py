#!/usr/bin/env python3
"""Generate the golden fixtures by running the reference CPU program.

The reference program is compiled from its own sources by ``make -C oracle ref``
into ``oracle/_ref/RTM3D.bin`` (never committed).  This script writes the
velocity model and input JSONs of a few small cases into a scratch directory,
runs the reference binary on them exactly as a user would
(``RTM3D.bin input.json``), and stores inputs and outputs as compressed
``.npz`` fixtures next to this file.  Only this container can run it (the
reference tree is needed to build the binary); the fixtures travel instead.

Cases (all grids non-cubic with distinct spacings, so a transposed axis fails):
  small_mod   3-layer model, ABC modeling, 2 shots, 2 time chunks, order 8
  small_hbc   HBC migration of small_mod's records (nthreads = 1, vpe saved)
  small_rbc   RBC migration of the same records (nthreads = 1, vpe saved)
  o4_mod      order-4 modeling, blen 6, odd sizes
  small_fmig  frequency-domain migration (fmig) of small_mod's records,
              1-30 Hz (7 frequencies), nthreads = 1, vpe saved

``make_golden.py fmig`` regenerates only small_fmig (from small_mod.npz's
records), leaving the other fixtures untouched.

``make_golden.py c0`` runs the 128^3 three-layer modeling case in full (1000
steps, 64 x 64 receivers, border 32; about two minutes on 8 cores) and stores
what pins the whole record without shipping its 16 MB: the SHA-256 of the
reference's seismogram file, every trace's float64 L2 norm, and a decimated
sub-record (every 8th receiver in x and y, every 4th sample) in
``c0_modeling.npz``.
"""
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REFBIN = os.path.join(REPO, "oracle", "_ref", "RTM3D.bin")


def three_layer(nx, ny, nz, z1, z2, v=(1500.0, 2500.0, 3500.0)):
    m = np.empty((nx, ny, nz), np.float32)
    m[:, :, :z1] = v[0]
    m[:, :, z1:z2] = v[1]
    m[:, :, z2:] = v[2]
    return m


def base_params(name, nx, ny, nz, blen, nt, ntstep, dx, dy, dz, order):
    vmax = 3500.0
    dt = float(np.float32(0.4 * min(dx, dy, dz) / vmax))
    return {
        "mname": name, "vpfile": "vel/model.bin", "vpefile": "vel/model_ext.bin",
        "save_vpe_file": False, "load_vpe_from_file": False,
        "outdir": "out", "tmpdir": "tmp", "datdir": "dat",
        "nx": nx, "ny": ny, "nz": nz, "blen": blen,
        "width_m": int(nx * dx), "height_m": int(ny * dy), "depth_m": int(nz * dz),
        "nt": nt, "dz": dz, "dy": dy, "dx": dx, "dt": dt,
        "fpeak": 15.0, "fmax": 30.0, "taper_factor": 0.9, "stencil_order": order,
        "modeling": True, "migration": False, "rtm_type": "acoustic",
        "boundary_condition": "hbc",
        "source_total": 1, "source_start_x": nx // 2, "source_start_y": ny // 2,
        "source_depth_z": 3, "source_distance_x": 1, "source_distance_y": 1,
        "source_count_x": 1, "source_count_y": 1,
        "receiver_total": 1, "receiver_start_x": 0, "receiver_start_y": 0,
        "receiver_depth_z": 1, "receiver_distance_x": 1, "receiver_distance_y": 1,
        "receiver_count_x": nx, "receiver_count_y": ny,
        "ntstep": ntstep, "nthreads": 8, "save_snapshots": False,
        "distributed_grid": False, "snapshot_step": 10, "fmig": False,
        "fmig_max_freq": 30.0, "fmig_min_freq": 1.0, "fmig_nwstep": 1,
        "fmig_distributed_imaging": False, "fpga_xclbin": "", "fpga_fwd_kernel": "",
        "fpga_bwd_kernel": "", "verbose": False,
    }


def run(workdir, params):
    for d in ("vel", "out", "tmp", "dat"):
        os.makedirs(os.path.join(workdir, d), exist_ok=True)
    path = os.path.join(workdir, "input.json")
    with open(path, "w") as f:
        json.dump(params, f, indent=1)
    r = subprocess.run([REFBIN, path], cwd=workdir, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout[-4000:] + r.stderr[-4000:])
        raise SystemExit("reference run failed")


def shots(p):
    out = []
    for ix in range(p["source_count_x"]):
        for iy in range(p["source_count_y"]):
            out.append((ix * p["source_distance_x"] + p["source_start_x"],
                        iy * p["source_distance_y"] + p["source_start_y"],
                        p["source_depth_z"]))
    return out


def seism_name(p, s, itstep):
    return "RTMSEISMOGRAM_S%dx%dx%d_NXNY_%dx%d_NTSTART%d_NTSTEP%d.seism" % (
        s[0], s[1], s[2], p["receiver_count_x"], p["receiver_count_y"], itstep, p["ntstep"])


def read_records(workdir, p):
    """Per shot: the ntstep chunks concatenated, [nchunk][nrx][nry][ntstep]."""
    recs = []
    for s in shots(p):
        chunks = []
        for itstep in range(0, p["nt"], p["ntstep"]):
            a = np.fromfile(os.path.join(workdir, "dat", seism_name(p, s, itstep)), np.float32)
            chunks.append(a.reshape(p["receiver_count_x"], p["receiver_count_y"], p["ntstep"]))
        recs.append(np.stack(chunks))
    return np.stack(recs)


def image_name(p):
    return "RTMIMG_%s_NS%d_NXNYNZ_%dx%dx%d_NT%d.bin.NP1" % (
        p["mname"].upper(), p["source_count_x"] * p["source_count_y"], p["nx"], p["ny"],
        p["nz"], p["nt"])


def small_params():
    nx, ny, nz, blen = 36, 44, 40, 10
    p = base_params("small", nx, ny, nz, blen, 300, 150, 10.0, 12.0, 8.0, 8)
    p.update({"source_count_x": 2, "source_distance_x": 5, "source_start_x": 17,
              "source_start_y": 21, "source_depth_z": 3,
              "receiver_start_x": 1, "receiver_start_y": 2, "receiver_depth_z": 1,
              "receiver_distance_x": 2, "receiver_distance_y": 3,
              "receiver_count_x": 16, "receiver_count_y": 13})
    return p


def fmig_case(wd, p, recs):
    """Frequency-domain migration of the small case's records."""
    for k, s in enumerate(shots(p)):
        for c, itstep in enumerate(range(0, p["nt"], p["ntstep"])):
            recs[k][c].astype(np.float32).tofile(os.path.join(wd, "dat", seism_name(p, s, itstep)))
    q = dict(p)
    q.update({"modeling": False, "migration": True, "boundary_condition": "hbc", "fmig": True,
              "fmig_min_freq": 1.0, "fmig_max_freq": 30.0, "fmig_nwstep": 1,
              "nthreads": 1, "save_vpe_file": True, "mname": "small_fmig"})
    run(wd, q)
    img = np.fromfile(os.path.join(wd, "out", image_name(q)), np.float32)
    img = img.reshape(p["nx"], p["ny"], p["nz"])
    vpe = np.fromfile(os.path.join(wd, "vel", "model_ext.bin"), np.float32)
    b = p["blen"]
    vpe = vpe.reshape(p["nx"] + 2 * b, p["ny"] + 2 * b, p["nz"] + 2 * b)
    np.savez_compressed(os.path.join(HERE, "small_fmig.npz"), image=img, vpe=vpe)
    return q


def main_fmig():
    tmp = tempfile.mkdtemp(prefix="rtmgold_")
    try:
        p = small_params()
        fx = np.load(os.path.join(HERE, "small_mod.npz"), allow_pickle=False)
        wd = os.path.join(tmp, "small")
        for d in ("vel", "out", "tmp", "dat"):
            os.makedirs(os.path.join(wd, d), exist_ok=True)
        fx["vel"].tofile(os.path.join(wd, "vel", "model.bin"))
        q = fmig_case(wd, p, fx["records"])
    finally:
        shutil.rmtree(tmp)
    path = os.path.join(HERE, "cases.json")
    with open(path) as f:
        cases = json.load(f)
    cases["small_fmig"] = q
    with open(path, "w") as f:
        json.dump(cases, f, indent=1, sort_keys=True)
    print("fixture written: small_fmig")


def c0_params():
    """The 128^3 three-layer modeling case (15 Hz, dt from CFL 0.4, taper 0.9)."""
    nx = ny = nz = 128
    d = 10.0
    p = base_params("l3_128", nx, ny, nz, 32, 1000, 1000, d, d, d, 8)
    p.update({"source_start_x": 64, "source_start_y": 64, "source_depth_z": 2,
              "receiver_start_x": 0, "receiver_start_y": 0, "receiver_depth_z": 0,
              "receiver_distance_x": 2, "receiver_distance_y": 2,
              "receiver_count_x": 64, "receiver_count_y": 64,
              "receiver_total": 4096, "fmax": 30.0})
    return p


def main_c0():
    import hashlib

    tmp = tempfile.mkdtemp(prefix="rtmgold_")
    try:
        p = c0_params()
        wd = os.path.join(tmp, "c0")
        os.makedirs(os.path.join(wd, "vel"))
        three_layer(128, 128, 128, 40, 80).tofile(os.path.join(wd, "vel", "model.bin"))
        run(wd, p)
        path = os.path.join(wd, "dat", seism_name(p, shots(p)[0], 0))
        raw = open(path, "rb").read()
        rec = np.frombuffer(raw, np.float32).reshape(64, 64, 1000)
        np.savez_compressed(os.path.join(HERE, "c0_modeling.npz"),
                            sha256=np.frombuffer(hashlib.sha256(raw).digest(), np.uint8),
                            trace_l2=np.sqrt((rec.astype(np.float64) ** 2).sum(axis=2)),
                            sub=np.ascontiguousarray(rec[::8, ::8, ::4]))
    finally:
        shutil.rmtree(tmp)
    path = os.path.join(HERE, "cases.json")
    with open(path) as f:
        cases = json.load(f)
    cases["c0_modeling"] = p
    with open(path, "w") as f:
        json.dump(cases, f, indent=1, sort_keys=True)
    print("fixture written: c0_modeling")


def main():
    if not os.path.exists(REFBIN):
        raise SystemExit("build the reference first: make -C oracle ref")
    if sys.argv[1:] == ["fmig"]:
        return main_fmig()
    if sys.argv[1:] == ["c0"]:
        return main_c0()
    cases = {}
    tmp = tempfile.mkdtemp(prefix="rtmgold_")
    try:
        # ---- small: modeling, then HBC and RBC migration of its records ----
        nx, ny, nz, blen = 36, 44, 40, 10
        vel = three_layer(nx, ny, nz, 13, 27)
        p = small_params()
        wd = os.path.join(tmp, "small")
        os.makedirs(os.path.join(wd, "vel"))
        vel.tofile(os.path.join(wd, "vel", "model.bin"))
        run(wd, p)
        recs = read_records(wd, p)
        np.savez_compressed(os.path.join(HERE, "small_mod.npz"), vel=vel, records=recs)
        cases["small_mod"] = dict(p)
        for bc in ("hbc", "rbc"):
            q = dict(p)
            q.update({"modeling": False, "migration": True, "boundary_condition": bc,
                      "nthreads": 1, "save_vpe_file": True, "mname": "small_" + bc})
            run(wd, q)
            img = np.fromfile(os.path.join(wd, "out", image_name(q)), np.float32)
            img = img.reshape(nx, ny, nz)
            vpe = np.fromfile(os.path.join(wd, "vel", "model_ext.bin"), np.float32)
            vpe = vpe.reshape(nx + 2 * blen, ny + 2 * blen, nz + 2 * blen)
            np.savez_compressed(os.path.join(HERE, "small_%s.npz" % bc), image=img, vpe=vpe)
            cases["small_" + bc] = q
        cases["small_fmig"] = fmig_case(wd, p, recs)

        # ---- order-4 modeling, odd sizes, one chunk ----
        nx, ny, nz, blen = 21, 27, 25, 6
        vel = three_layer(nx, ny, nz, 8, 16)
        p = base_params("o4", nx, ny, nz, blen, 160, 160, 9.0, 11.0, 10.0, 4)
        p.update({"source_start_x": 10, "source_start_y": 13, "source_depth_z": 2,
                  "receiver_start_x": 0, "receiver_start_y": 1, "receiver_depth_z": 0,
                  "receiver_distance_x": 1, "receiver_distance_y": 2,
                  "receiver_count_x": 21, "receiver_count_y": 13})
        wd = os.path.join(tmp, "o4")
        os.makedirs(os.path.join(wd, "vel"))
        vel.tofile(os.path.join(wd, "vel", "model.bin"))
        run(wd, p)
        np.savez_compressed(os.path.join(HERE, "o4_mod.npz"), vel=vel,
                            records=read_records(wd, p))
        cases["o4_mod"] = dict(p)
    finally:
        shutil.rmtree(tmp)
    with open(os.path.join(HERE, "cases.json"), "w") as f:
        json.dump(cases, f, indent=1, sort_keys=True)
    print("fixtures written:", ", ".join(sorted(cases)))


if __name__ == "__main__":
    main()
