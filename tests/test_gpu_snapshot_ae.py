"""Snapshot RTM with coded snapshots (rtmhip_snap_ae_forward_steps /
rtmhip_snap_ae_backward_steps) against the composition of calls it stands
for, bit for bit: the codes and scales of every snapshot equal
rtmhip_ae_bricks + rtmhip_ae_encode applied to the F32 snapshots of
rtmhip_snap_forward_steps (which the oracle pins), and the image equals
rtmhip_snap_backward_steps run on the snapshots that rtmhip_ae_decode
rebuilds from those codes.  The grid leaves partial bricks on every axis,
the border is odd, and both passes are split into chunks.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

B, NT = 11, 26
GRID = (36, 42, 70)                    # imaging region 36 x 42 x 48: 2 x 2 x 2 bricks
SRC = (17, 19, 2)
RCV = (1, 2, 3, 3, 4, 4, 5)


def _inputs(seed=5):
    rng = np.random.default_rng(seed)
    V = rng.uniform(0.0, 0.15, GRID).astype(np.float32)
    coefs = np.zeros(27, np.float32)
    c8 = np.array([-1 / 560, 8 / 315, -1 / 5, 8 / 5, -205 / 72, 8 / 5, -1 / 5, 8 / 315, -1 / 560])
    for d, h in enumerate((10.0, 11.0, 9.0)):
        coefs[d * 9:(d + 1) * 9] = (c8 / (h * h)).astype(np.float32)
    g = np.exp(-(np.sqrt(-np.log(0.9)) / B * np.arange(B)) ** 2).astype(np.float32)
    ricker = rng.uniform(-1, 1, NT).astype(np.float32)
    rcv = rng.uniform(-1, 1, (RCV[5], RCV[6], NT)).astype(np.float32)
    return V, coefs, g, ricker, rcv


@pytest.mark.parametrize("fusion", [-1, 0], ids=["tdefault", "tsep"])
@pytest.mark.parametrize("k", [1, 3])
def test_coded_snapshot_rtm_equals_composition(gpu, k, fusion):
    from rtm3d_demo_amd import ae
    print("coded snapshot rtm", k, fusion, flush=True)
    gpu.set_taper_fusion("f32", fusion)
    V, coefs, g, ricker, rcv = _inputs()
    lz = GRID[2] - 2 * B
    shot = gpu.make_shot(*GRID, B, 8, NT, NT, tuple(c + B for c in SRC), gpu.RcvGeom(*RCV))
    dV, dg, dsrc, drcv = (gpu.DeviceArray.from_host(a) for a in (V, g, ricker, rcv))
    codec = ae.Codec(max_bricks=3, seed=11)   # passes of 3 bricks: a ragged last pass
    cs = ae.CodedSnapshots(shot, k, codec)
    assert cs.nb == 8 and cs.nsnap == (NT - 1) // k + 1
    S1, S2 = gpu.FieldPair(GRID), gpu.FieldPair(GRID)
    snaps = gpu.snap_buffer(shot, "f32", k)
    R1, R2 = gpu.FieldPair(GRID), gpu.FieldPair(GRID)
    img1, img2 = gpu.DeviceArray(GRID).zero(), gpu.DeviceArray(GRID).zero()
    field = gpu.DeviceArray((GRID[0], GRID[1], lz))
    bricks = gpu.DeviceArray((cs.nb, ae.BRICK), np.uint16)
    scales = gpu.DeviceArray((cs.nb,))
    codes = gpu.DeviceArray((cs.nb, ae.CODES), np.uint16)
    region = (0, 0, 0, GRID[0], GRID[1], lz)
    try:
        for it0, n in ((0, 10), (10, NT - 10)):
            cs.forward_steps(dV, coefs, dg, dsrc, S1, it0, n)
            gpu.snap_forward_steps(shot, dV, coefs, dg, dsrc, S2, snaps, "f32", k, it0, n)
        got_codes, got_scales = cs.codes.download(), cs.scales.download()
        exact = snaps.download()
        decoded = np.empty_like(exact)
        for j in range(cs.nsnap):
            field.upload(exact[j])
            ae.cut_bricks(field, region, bricks, scales)
            codec.encode(bricks, cs.nb, codes)
            assert np.array_equal(got_scales[j].view(np.uint32), scales.download().view(np.uint32))
            assert np.array_equal(got_codes[j], codes.download()), j
            codes.upload(got_codes[j])
            scales.upload(got_scales[j])
            field.zero()
            codec.decode(codes, scales, cs.nb, field, region)
            decoded[j] = field.download()
        assert np.abs(decoded).max() > 0
        snaps.upload(decoded)
        for hi, n in ((NT - 1, 9), (NT - 10, NT - 9)):
            cs.backward_steps(dV, coefs, dg, R1, drcv, img1, hi, n)
            gpu.snap_backward_steps(shot, dV, coefs, dg, R2, snaps, "f32", k, drcv, img2, hi, n)
        got, want = img1.download(), img2.download()
    finally:
        gpu.set_taper_fusion("f32", -1)
        for d in (dV, dg, dsrc, drcv, S1, S2, snaps, R1, R2, img1, img2, field, bricks, scales,
                  codes):
            d.free()
        cs.free()
        codec.destroy()
    assert np.abs(want).max() > 0
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_coded_snapshots_reject_16bit_wavefields_and_bad_interval(gpu):
    from rtm3d_demo_amd import ae
    V, coefs, g, ricker, rcv = _inputs()
    dV, dg, dsrc = (gpu.DeviceArray.from_host(a) for a in (V, g, ricker))
    codec = ae.Codec(max_bricks=8, seed=1)
    shot = gpu.make_shot(*GRID, B, 8, NT, NT, tuple(c + B for c in SRC), gpu.RcvGeom(*RCV))
    cs = ae.CodedSnapshots(shot, 2, codec)
    shot16 = gpu.make_shot(*GRID, B, 8, NT, NT, tuple(c + B for c in SRC), gpu.RcvGeom(*RCV),
                           "f16")
    S16 = gpu.FieldPair(GRID, np.uint16)
    S = gpu.FieldPair(GRID)
    try:
        cs.shot = shot16
        with pytest.raises(gpu.RtmError) as e:
            cs.forward_steps(dV, coefs, dg, dsrc, S16, 0, 1)
        assert e.value.code == gpu.ERR_UNSUPPORTED
        cs.shot, cs.k = shot, 0
        with pytest.raises(gpu.RtmError, match="interval"):
            cs.forward_steps(dV, coefs, dg, dsrc, S, 0, 1)
        cs.k = 2
        with pytest.raises(gpu.RtmError, match="time range"):
            cs.forward_steps(dV, coefs, dg, dsrc, S, NT - 1, 2)
    finally:
        for d in (dV, dg, dsrc, S16, S):
            d.free()
        cs.free()
        codec.destroy()


def test_codec_properties_on_full_salt_region(gpu):
    """Encode / decode of the imaging region of the 592 x 592 x 336 salt grid
    (z in [40, 296): 2888 bricks), size-independent properties:
    * scaling the field by 2^k scales every decoded value by 2^k exactly
      (bricks are normalised by their largest magnitude, so the codes are the
      same and the scale carries the power of two);
    * all-zero bricks decode to zero;
    * decode writes the region's cells and no other."""
    from rtm3d_demo_amd import ae
    shape, b = (592, 592, 336), 40
    region = ae.region_of(shape, b)
    nb = ae.region_bricks(region)
    assert nb == 2888
    rng = np.random.default_rng(8)
    f = np.empty(shape, np.float32)
    for x in range(shape[0]):
        f[x] = rng.standard_normal(shape[1:], dtype=np.float32)
    f[:, :, 232:] = 0.0                      # the bricks at z >= 232 hold zeros only
    codec = ae.Codec(max_bricks=1024, seed=5)
    dF = gpu.DeviceArray.from_host(f)
    bricks = gpu.DeviceArray((nb, ae.BRICK), np.uint16)
    scales = gpu.DeviceArray((nb,))
    codes = gpu.DeviceArray((nb, ae.CODES), np.uint16)
    out = gpu.DeviceArray(shape)
    try:
        def roundtrip(field):
            dF.upload(field)
            ae.cut_bricks(dF, region, bricks, scales)
            codec.encode(bricks, nb, codes)
            gpu.check(gpu.load().rtmhip_fill(out.ptr, out.size, 7.0))
            codec.decode(codes, scales, nb, out, region)
            return out.download(), codes.download(), scales.download()

        d1, c1, s1 = roundtrip(f)
        x0, y0, z0, x1, y1, z1 = region
        inside = d1[x0:x1, y0:y1, z0:z1]
        assert np.isfinite(inside).all()
        assert (d1[:, :, :z0] == 7.0).all() and (d1[:, :, z1:] == 7.0).all()
        assert (inside[:, :, 232 - z0:] == 0.0).all()
        assert (s1.reshape(19, 19, 8)[:, :, 6:] == 0.0).all()
        assert (s1.reshape(19, 19, 8)[:, :, :6] > 0.0).all()
        for k in (-3, 5):
            dk, ck, sk = roundtrip(f * np.float32(2.0 ** k))
            assert np.array_equal(ck, c1)
            assert np.array_equal(sk, s1 * np.float32(2.0 ** k))
            assert np.array_equal(dk[x0:x1, y0:y1, z0:z1], inside * np.float32(2.0 ** k))
    finally:
        for d in (dF, bricks, scales, codes, out):
            d.free()
        codec.destroy()
