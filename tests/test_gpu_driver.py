"""The rtm3d driver binary end to end on the GPU: the reference's JSON in,
the reference's files out, bit-identical to the reference program's files
(seismogram chunks, extended velocity, migrated image)."""
import json
import os
import subprocess

import numpy as np
import pytest

from rtm3d_demo_amd import DRIVER

pytestmark = pytest.mark.gpu


def _run(wd, params):
    path = os.path.join(wd, "input.json")
    with open(path, "w") as f:
        json.dump(params, f)
    r = subprocess.run([DRIVER, path], cwd=wd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    return r


def _seism(p, s, itstep):
    return "RTMSEISMOGRAM_S%dx%dx%d_NXNY_%dx%d_NTSTART%d_NTSTEP%d.seism" % (
        s[0], s[1], s[2], p["receiver_count_x"], p["receiver_count_y"], itstep, p["ntstep"])


def _shots(p):
    return [(ix * p["source_distance_x"] + p["source_start_x"],
             iy * p["source_distance_y"] + p["source_start_y"], p["source_depth_z"])
            for ix in range(p["source_count_x"]) for iy in range(p["source_count_y"])]


@pytest.fixture
def workdir(tmp_path, golden):
    for d in ("vel", "out", "tmp", "dat"):
        (tmp_path / d).mkdir()
    golden("small_mod")["vel"].tofile(str(tmp_path / "vel" / "model.bin"))
    return str(tmp_path)


def test_driver_modeling_then_migration(workdir, cases, golden):
    print("driver small case", flush=True)
    p = dict(cases["small_mod"])
    p["verbose"] = True
    _run(workdir, p)
    recs = golden("small_mod")["records"]
    for k, s in enumerate(_shots(p)):
        for c, itstep in enumerate(range(0, p["nt"], p["ntstep"])):
            got = np.fromfile(os.path.join(workdir, "dat", _seism(p, s, itstep)), np.float32)
            assert np.array_equal(got.view(np.uint32), recs[k, c].ravel().view(np.uint32))
    for bc in ("hbc", "rbc"):
        q = dict(cases["small_" + bc])
        _run(workdir, q)
        vpe = np.fromfile(os.path.join(workdir, "vel", "model_ext.bin"), np.float32)
        assert np.array_equal(vpe, golden("small_" + bc)["vpe"].ravel())
        name = "RTMIMG_%s_NS2_NXNYNZ_%dx%dx%d_NT%d.bin.NP1" % (
            q["mname"].upper(), q["nx"], q["ny"], q["nz"], q["nt"])
        img = np.fromfile(os.path.join(workdir, "out", name), np.float32)
        assert np.array_equal(img.view(np.uint32), golden("small_" + bc)["image"].ravel().view(np.uint32))


def test_driver_snapshots(workdir, cases):
    p = dict(cases["small_mod"])
    p.update({"save_snapshots": True, "snapshot_step": 50, "source_count_x": 1})
    _run(workdir, p)
    nye, nze = p["ny"] + 2 * p["blen"], p["nz"] + 2 * p["blen"]
    name = "RTMSNAPSHOT_S%dx%d_%dx%dx%d.snapshot.RTM_MOD" % (
        p["source_start_x"] + p["blen"], p["source_start_y"] + p["blen"], p["nt"] // 50, nye, nze)
    snap = np.fromfile(os.path.join(workdir, "tmp", name), np.float32)
    assert snap.size == (p["nt"] // 50) * nye * nze
    assert np.isfinite(snap).all() and np.abs(snap[-nye * nze:]).max() > 0


def test_driver_reports_dispersion(workdir, cases):
    """An unstable time step blows the field up to NaN: exit status 1 with
    the reference's 'Numerical Dispersion' message."""
    p = dict(cases["small_mod"])
    p.update({"dt": 0.05, "nt": 300, "ntstep": 300, "source_count_x": 1})
    path = os.path.join(workdir, "input.json")
    with open(path, "w") as f:
        json.dump(p, f)
    r = subprocess.run([DRIVER, path], cwd=workdir, capture_output=True, text=True, timeout=300)
    assert r.returncode == 1 and "Numerical Dispersion" in r.stdout


@pytest.mark.parametrize("fmt, mode", [("f32", "2"), ("f16", "1"), ("bf16", "0")])
def test_driver_wave_format_and_backward_mode(workdir, cases, golden, oracle, fmt, mode):
    """RTM3D_WAVE_FORMAT / RTM3D_BACKWARD_MODE: modeling records and the HBC
    image written by the driver equal the oracle's pipeline in that storage
    format (the backward mode never changes results)."""
    print("driver", fmt, "mode", mode, flush=True)
    env = dict(os.environ, RTM3D_WAVE_FORMAT=fmt, RTM3D_BACKWARD_MODE=mode)
    p = dict(cases["small_mod"])
    path = os.path.join(workdir, "input.json")
    with open(path, "w") as f:
        json.dump(p, f)
    r = subprocess.run([DRIVER, path], cwd=workdir, capture_output=True, text=True,
                       timeout=300, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    vel = golden("small_mod")["vel"]
    ofmt = None if fmt == "f32" else fmt
    recs = oracle.pipeline_modeling(p, vel, np.float32, ofmt)
    s0 = _shots(p)[0]
    got = np.fromfile(os.path.join(workdir, "dat", _seism(p, s0, 0)), np.float32)
    assert np.array_equal(got.view(np.uint32), recs[0, 0].ravel().view(np.uint32))
    q = dict(cases["small_hbc"], load_vpe_from_file=False, save_vpe_file=True, nthreads=1)
    with open(path, "w") as f:
        json.dump(q, f)
    r = subprocess.run([DRIVER, path], cwd=workdir, capture_output=True, text=True,
                       timeout=300, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    vpe = np.fromfile(os.path.join(workdir, "vel", "model_ext.bin"), np.float32)
    vpe = vpe.reshape(golden("small_hbc")["vpe"].shape)
    name = "RTMIMG_%s_NS2_NXNYNZ_%dx%dx%d_NT%d.bin.NP1" % (
        q["mname"].upper(), q["nx"], q["ny"], q["nz"], q["nt"])
    img = np.fromfile(os.path.join(workdir, "out", name), np.float32)
    want = oracle.pipeline_migration(q, vpe, recs, np.float32, ofmt)
    assert np.array_equal(img.view(np.uint32), want.ravel().view(np.uint32))


@pytest.mark.parametrize("fmt", ["f32", "f16"])
def test_driver_upb_spill(workdir, cases, golden, oracle, fmt):
    """RTM3D_UPB_SPILL=1: one receiver chunk of the HBC upper-border history
    on the device, the chunks in host memory between the passes (the
    reference spills them to files); the image is unchanged: the reference's
    for FP32, the oracle's FP16 pipeline for FP16."""
    print("driver upb spill", fmt, flush=True)
    env = dict(os.environ, RTM3D_WAVE_FORMAT=fmt, RTM3D_UPB_SPILL="1")
    p = dict(cases["small_mod"])
    path = os.path.join(workdir, "input.json")
    with open(path, "w") as f:
        json.dump(p, f)
    r = subprocess.run([DRIVER, path], cwd=workdir, capture_output=True, text=True,
                       timeout=300, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    q = dict(cases["small_hbc"])
    assert q["nt"] // q["ntstep"] >= 2
    with open(path, "w") as f:
        json.dump(q, f)
    r = subprocess.run([DRIVER, path], cwd=workdir, capture_output=True, text=True,
                       timeout=300, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    name = "RTMIMG_%s_NS2_NXNYNZ_%dx%dx%d_NT%d.bin.NP1" % (
        q["mname"].upper(), q["nx"], q["ny"], q["nz"], q["nt"])
    img = np.fromfile(os.path.join(workdir, "out", name), np.float32)
    if fmt == "f32":
        want = golden("small_hbc")["image"]
    else:
        vel = golden("small_mod")["vel"]
        recs = oracle.pipeline_modeling(p, vel, np.float32, fmt)
        want = oracle.pipeline_migration(q, golden("small_hbc")["vpe"], recs, np.float32, fmt)
    assert np.array_equal(img.view(np.uint32), want.ravel().view(np.uint32))
    bad = dict(os.environ, RTM3D_UPB_SPILL="2")
    r = subprocess.run([DRIVER, path], cwd=workdir, capture_output=True, text=True,
                       timeout=300, env=bad)
    assert r.returncode == 1 and "RTM3D_UPB_SPILL" in r.stdout + r.stderr


def test_driver_fmig(workdir, cases, golden):
    """Frequency-domain migration (fmig) through the driver: the reference's
    image bit for bit, with its running image across the two shots."""
    q = dict(cases["small_fmig"])
    recs = golden("small_mod")["records"]
    for k, s in enumerate(_shots(q)):
        for c, itstep in enumerate(range(0, q["nt"], q["ntstep"])):
            recs[k][c].astype(np.float32).tofile(os.path.join(workdir, "dat", _seism(q, s, itstep)))
    _run(workdir, q)
    vpe = np.fromfile(os.path.join(workdir, "vel", "model_ext.bin"), np.float32)
    assert np.array_equal(vpe, golden("small_fmig")["vpe"].ravel())
    name = "RTMIMG_%s_NS2_NXNYNZ_%dx%dx%d_NT%d.bin.NP1" % (
        q["mname"].upper(), q["nx"], q["ny"], q["nz"], q["nt"])
    img = np.fromfile(os.path.join(workdir, "out", name), np.float32)
    assert np.array_equal(img.view(np.uint32), golden("small_fmig")["image"].ravel().view(np.uint32))
