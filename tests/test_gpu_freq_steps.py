"""The frequency-domain migration loop (rtmhip_freq_steps) against the
reference's operator sequence per step on the oracle
(RTMFiniteDifferencesKernel_FreqDomain.cpp:300-357): both fields stepped,
taper3D of the four grids, source at st, receivers of lt set, contributions
and image of the frequency group at st, swaps -- bit for bit on non-zero
fields, over two receiver chunks, in whole-chunk calls and split calls.

The default taper form applies the tapers to the new fields' output (even
nz: two-z waves, odd nz: one-z waves); a receiver line that reaches the x
border falls back to the in-place passes, as does taper fusion mode 0.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NT, NTSTEP, B = 12, 6, 6


@pytest.mark.parametrize("shape", [(40, 44, 36), (30, 34, 41)])
@pytest.mark.parametrize("kind, fusion", [("inside", -1), ("border", -1), ("inside", 0)],
                         ids=["inside", "border", "separate"])
@pytest.mark.parametrize("splits", [(6, 6), (4, 2, 6)])
def test_freq_steps_match_operator_sequence(gpu, oracle, shape, kind, fusion, splits):
    # receivers from x = B + 1 every 3 cells: "border" reaches x >= nx - B
    # (still inside the grid), "inside" stays clear of every border
    cx = -(-(shape[0] - B - 7) // 3) + 1 if kind == "border" else 3
    rgeom = (1, 2, 3, 3, 2, cx, 4)
    last = B + 1 + 3 * (cx - 1)
    assert last < shape[0] and (last >= shape[0] - B) == (kind == "border")
    print("freq steps", shape, rgeom, fusion, splits, flush=True)
    order, nwstep, iw0 = 8, 2, 0
    rng = np.random.default_rng(51)
    Ps, PPs, Pr, PPr, img = (rng.uniform(-1, 1, shape).astype(np.float32) for _ in range(5))
    v2 = rng.uniform(0.0, 0.15, shape).astype(np.float32)
    g = oracle.taper_table(B, np.float32(0.3))
    coefs = oracle.stencil_coefs(order, 10.0, 11.0, 9.0)
    ricker = rng.uniform(-1, 1, NT).astype(np.float32)
    rcv = rng.uniform(-1, 1, (rgeom[5], rgeom[6], NT)).astype(np.float32)
    inner = tuple(n - 2 * B for n in shape)
    acc = [rng.uniform(-1, 1, (nwstep,) + inner).astype(np.float32) for _ in range(4)]
    w2, kre, kim = oracle.freq_tables(NT, np.float32(0.004), np.float32(2.0), np.float32(40.0))
    nw = min(3, len(w2))
    w2, kre, kim = (np.ascontiguousarray(a[:nw]) for a in (w2, kre, kim))
    src = (shape[0] // 2, shape[1] // 2, B + 2)
    # the device run: chunk 1 (steps 11..6), then chunk 0 (5..0)
    gpu.set_taper_fusion("f32", fusion)
    S, R = gpu.FieldPair(shape), gpu.FieldPair(shape)
    S.a.upload(Ps)
    S.b.upload(PPs)
    R.a.upload(Pr)
    R.b.upload(PPr)
    d = {k: gpu.DeviceArray.from_host(a) for k, a in
         dict(V=v2, g=g, ricker=ricker, img=img, kre=kre, kim=kim, w2=w2).items()}
    dacc = [gpu.DeviceArray.from_host(a) for a in acc]
    shot = gpu.make_shot(*shape, B, order, NT, NTSTEP, src, gpu.RcvGeom(*rgeom))
    try:
        hi = NT - 1
        for n in splits:
            chunk = (hi // NTSTEP) * NTSTEP
            dr = gpu.DeviceArray.from_host(np.ascontiguousarray(rcv[:, :, chunk:chunk + NTSTEP]))
            gpu.freq_steps(shot, d["V"], coefs, d["g"], d["ricker"], S, R, dr, d["img"], dacc,
                           d["kre"], d["kim"], d["w2"], nwstep, iw0, nw, hi, n)
            dr.free()
            hi -= n
        got = [S.P.download(), S.PP.download(), R.P.download(), R.PP.download(),
               d["img"].download()] + [a.download() for a in dacc]
    finally:
        gpu.set_taper_fusion("f32", -1)
        for x in list(d.values()) + dacc + [S, R]:
            x.free()
    og = oracle.RcvGeom(*rgeom)
    for rt in range(NT - 1, -1, -1):
        st, lt = NT - 1 - rt, rt % NTSTEP
        chunk = (rt // NTSTEP) * NTSTEP
        rc = np.ascontiguousarray(rcv[:, :, chunk:chunk + NTSTEP])
        oracle.step(Ps, PPs, v2, coefs, order)
        oracle.step(Pr, PPr, v2, coefs, order)
        for f in (Ps, PPs, Pr, PPr):
            oracle.taper(f, g, B, False)
        oracle.apply_source(PPs, *src, ricker[st])
        oracle.receivers(PPr, rc, B, og, NTSTEP, lt, True)
        oracle.freq_image(img, PPs, PPr, acc, kre, kim, w2, B, nwstep, iw0, NT, st)
        Ps, PPs = PPs, Ps
        Pr, PPr = PPr, Pr
    want = [Ps, PPs, Pr, PPr, img] + acc
    for k, (a, w) in enumerate(zip(got, want)):
        assert np.array_equal(a.view(np.uint32), w.view(np.uint32)), k


def test_freq_steps_reject_a_call_across_receiver_chunks(gpu, oracle):
    shape = (40, 44, 36)
    S, R = gpu.FieldPair(shape), gpu.FieldPair(shape)
    inner = tuple(n - 2 * B for n in shape)
    w2, kre, kim = oracle.freq_tables(NT, np.float32(0.004), np.float32(2.0), np.float32(40.0))
    d = [gpu.DeviceArray.from_host(np.ascontiguousarray(a)) for a in (kre, kim, w2)]
    acc = [gpu.DeviceArray((1,) + inner) for _ in range(4)]
    V, g, ric = gpu.DeviceArray(shape), gpu.DeviceArray((B,)), gpu.DeviceArray((NT,))
    rcv, img = gpu.DeviceArray((3, 4, NTSTEP)), gpu.DeviceArray(shape)
    shot = gpu.make_shot(*shape, B, 8, NT, NTSTEP, (20, 22, B + 2),
                         gpu.RcvGeom(1, 2, 3, 3, 2, 3, 4))
    coefs = oracle.stencil_coefs(8, 10.0, 11.0, 9.0)
    try:
        with pytest.raises(Exception, match="receiver chunks"):
            gpu.freq_steps(shot, V, coefs, g, ric, S, R, rcv, img, acc, d[0], d[1], d[2], 1, 0,
                           len(w2), 7, 4)
    finally:
        for x in [S, R, V, g, ric, rcv, img] + acc + d:
            x.free()
