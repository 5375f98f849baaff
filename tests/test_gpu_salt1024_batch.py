"""All 64 shot windows of the batched workload through ONE batch, checked
against the oracle's HBC shot.

The workload: 64 shots with uneven apertures (nx = ny from 256 to 448, nz from
256 to 320, multiples of 32) cut around their sources from the 1024 x 1024 x
320 salt model, HBC border 40, order 8.  The batch (rtmhip_batch_*) advances
all 64 shots in lockstep; its graphs hold 2 steps so that a short run replays
them in both directions (forward: 1 eager step, 3 replays, 1 eager step;
backward: the 2 seeding steps, 1 eager step, 2 replays, 1 eager step).

The oracle runs each shot alone with the reference's operator sequence
(orc_migrate_shot: rtm-3d/src/kernel/RTMFiniteDifferencesKernel_HBC.cpp:140-185
forward and :261-354 backward, the shot loop of
rtm-3d/src/controller/RTMAcousticController.cpp:188-200).  Every shot's last
forward wavefield and its image must be bit-identical.  Source samples and
records are seeded uniform noise (the batch path does not depend on where
they come from), so every operator contributes from the first step.

The batch runs once (module fixture: about 130 GB of device memory); the
comparison is split into groups of 8 shots so that each test finishes in well
under a minute.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NSHOTS, NT, G, GROUP = 64, 8, 2, 8
BLEN, ORDER = 40, 8


def _setup(host, workloads, model, shot):
    x0, y0, nxy, nz, src = shot
    p = workloads.salt_1024_params(nxy, nz, src, nt=NT)
    vel = np.ascontiguousarray(model[x0:x0 + nxy, y0:y0 + nxy, :nz])
    ve = host.extend_velocity(vel, BLEN, "hbc")
    return p, host.v2dt2(ve, np.float32(p["dt"]))


def _geom(mod, p):
    return mod.RcvGeom(p["receiver_start_x"], p["receiver_start_y"], p["receiver_depth_z"],
                       p["receiver_distance_x"], p["receiver_distance_y"],
                       p["receiver_count_x"], p["receiver_count_y"])


def _noise(k, p):
    rng = np.random.default_rng(1000 + k)
    ricker = rng.uniform(-1, 1, NT).astype(np.float32)
    rcv = rng.uniform(-1, 1, (p["receiver_count_x"], p["receiver_count_y"], NT)).astype(
        np.float32)
    return ricker, rcv


@pytest.fixture(scope="module")
def salt_batch(gpu):
    from rtm3d_demo_amd import host, workloads

    print("salt1024 batch: building 64 windows", flush=True)
    model = workloads.salt_1024_model()
    shots = workloads.salt_1024_shots(NSHOTS)
    d = 12.5
    coefs = host.stencil_coefs(ORDER, d, d, d)
    g = host.taper_table(BLEN, np.float32(0.9))
    dg = gpu.DeviceArray.from_host(g)
    members, saved = [], []
    try:
        for k, s in enumerate(shots):
            p, v2 = _setup(host, workloads, model, s)
            ricker, rcv = _noise(k, p)
            shape = v2.shape
            src = tuple(c + BLEN for c in s[4])
            m = dict(shot=gpu.make_shot(*shape, BLEN, ORDER, NT, NT, src, _geom(gpu, p)),
                     V=gpu.DeviceArray.from_host(v2), ricker=gpu.DeviceArray.from_host(ricker),
                     rcv=gpu.DeviceArray.from_host(rcv), img=gpu.DeviceArray(shape).zero(),
                     upb=gpu.DeviceArray((NT, shape[0], shape[1], ORDER // 2)).zero(),
                     S=gpu.FieldPair(shape), R=gpu.FieldPair(shape))
            members.append(m)
        print("salt1024 batch: running %d shots x %d steps" % (NSHOTS, NT), flush=True)
        batch = gpu.Batch(members, coefs, dg)
        try:
            batch.set_graph(G)
            batch.forward_steps(0, NT)
            for m in members:
                keep = gpu.DeviceArray(m["S"].PP.shape)
                gpu.check(gpu.load().rtmhip_copy_device(keep.ptr, m["S"].PP.ptr, keep.nbytes))
                saved.append(keep)
            batch.backward_steps(NT - 1, NT)
            gpu.synchronize()
        finally:
            batch.destroy()
        yield shots, members, saved, coefs, g, model
    finally:
        for m in members:
            for key in ("V", "ricker", "rcv", "img", "upb", "S", "R"):
                m[key].free()
        for a in saved:
            a.free()
        dg.free()


@pytest.mark.parametrize("group", range(NSHOTS // GROUP))
def test_salt1024_batch_matches_oracle(salt_batch, oracle, group):
    from rtm3d_demo_amd import host, workloads

    shots, members, saved, coefs, g, model = salt_batch
    print("salt1024 batch: oracle shots %d-%d" % (group * GROUP, group * GROUP + GROUP - 1),
          flush=True)
    for k in range(group * GROUP, (group + 1) * GROUP):
        x0, y0, nxy, nz, src = shots[k]
        p, v2 = _setup(host, workloads, model, shots[k])
        ricker, rcv = _noise(k, p)
        last = np.zeros(v2.shape, np.float32)
        want = oracle.migrate_shot(v2, coefs, g, ricker, BLEN, ORDER, NT, NT, src,
                                   _geom(oracle, p), rcv, True, fwd_last=last)
        got_last = saved[k].download()
        assert np.array_equal(got_last.view(np.uint32), last.view(np.uint32)), \
            "shot %d (%dx%dx%d): last forward wavefield differs" % (k, nxy, nxy, nz)
        got = members[k]["img"].download()
        assert np.abs(want).max() > 0
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), \
            "shot %d (%dx%dx%d): image differs" % (k, nxy, nxy, nz)
