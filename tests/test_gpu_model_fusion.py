"""Modeling steps (rtmhip_model_steps) with taper3D of both fields fused into
the stencil (taper fusion mode 3 and the default: applied to each step's
output; mode 2: applied on load), or applied by the separate in-place pass
(mode 0),
against the reference's operator sequence on the oracle (swap, taper3D(P),
taper3D(PP), step, source, save receivers), bit for bit, for every wavefield
format.

The output form needs the source and the receivers off the borders; the
wide receiver line reaches the x border of the narrower grids, where the
default falls back to the separate pass, and the inner line keeps the output
form on every grid.  Even nz runs it on two-z waves, odd nz on one-z waves.

The fields start non-zero everywhere, so the front-right edge of taper3D
(cells x < blen, y >= ny-blen, z inside take the face-tapered value of
(blen, y, z)) and the frame the stencil does not update both show in the
result.  Calls are split so that the pending taper between calls and the
first step's single PP taper are exercised; an odd border puts the taper
edge at an odd z; fixed short x chunks exercise the launcher's minimum
chunk (the first chunk must reach plane blen).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GRIDS = [(40, 44, 90), (32, 36, 340), (30, 34, 91)]
RCV = (1, 2, 3, 3, 4, 4, 5)      # start x, y, depth, dist x, y, count x, y
RCV_IN = (1, 2, 3, 2, 2, 3, 4)   # inside the borders of every grid below


@pytest.fixture(params=[((0, 0, 0), 0), ((64, 1, 8), 3), ((128, 1, 4), 0)],
                ids=lambda p: "t%d.%d.%d-x%d" % (*p[0], p[1]))
def tiling(request, gpu):
    t, xc = request.param
    for cls in range(3):
        gpu.set_tiling(cls, *t, xc)
    yield request.param
    for cls in range(3):
        gpu.set_tiling(cls, 0)


@pytest.mark.parametrize("fusion, rcvg", [(3, RCV_IN), (-1, RCV), (2, RCV), (0, RCV)],
                         ids=["tout-rin", "tdefault", "tload", "tsep"])
@pytest.mark.parametrize("splits", [(5,), (2, 1, 2)])
@pytest.mark.parametrize("b", [12, 11])
@pytest.mark.parametrize("shape", GRIDS)
@pytest.mark.parametrize("fmt", ["f32", "f16", "bf16"])
def test_model_steps_match_operator_sequence(gpu, oracle, fmt, shape, b, splits, fusion, rcvg,
                                             tiling):
    print("model fusion", fmt, shape, b, splits, fusion, rcvg, tiling, flush=True)
    order, nt = 8, 5
    rng = np.random.default_rng(41)
    conv = (lambda a: a) if fmt == "f32" else (lambda a: oracle.to_format(a, fmt))
    ofmt = () if fmt == "f32" else (fmt,)
    P0 = conv(rng.uniform(-1, 1, shape).astype(np.float32))
    PP0 = conv(rng.uniform(-1, 1, shape).astype(np.float32))
    v2 = rng.uniform(0.0, 0.15, shape).astype(np.float32)
    g = oracle.taper_table(b, np.float32(0.3))
    coefs = oracle.stencil_coefs(order, 10.0, 11.0, 9.0)
    src = rng.uniform(-1, 1, nt).astype(np.float32)
    sxe = (shape[0] // 2, shape[1] // 2, b + 2)
    ogeom = oracle.RcvGeom(*rcvg)
    P, PP = P0.copy(), PP0.copy()
    rec = np.zeros((rcvg[5], rcvg[6], nt), np.float32)
    for it in range(nt):
        P, PP = PP, P
        oracle.taper(P, g, b, False, *ofmt)
        oracle.taper(PP, g, b, False, *ofmt)
        oracle.step(P, PP, v2, coefs, order, *ofmt)
        oracle.apply_source(PP, *sxe, src[it], *ofmt)
        oracle.receivers(PP, rec, b, ogeom, nt, it, False, *ofmt)
    gpu.set_taper_fusion(fmt, fusion)
    wd = gpu.WAVE_FORMATS[fmt][1]
    F = gpu.FieldPair(shape, wd)
    F.a.upload(P0)
    F.b.upload(PP0)
    V, dg, ds = (gpu.DeviceArray.from_host(a) for a in (v2, g, src))
    drec = gpu.DeviceArray(rec.shape).zero()
    shot = gpu.make_shot(*shape, b, order, nt, nt, sxe, gpu.RcvGeom(*rcvg), fmt)
    try:
        it0 = 0
        for n in splits:
            gpu.model_steps(shot, V, coefs, dg, ds, F, drec, it0, n)
            it0 += n
        gotP, gotPP, got_rec = F.P.download(), F.PP.download(), drec.download()
    finally:
        gpu.set_taper_fusion(fmt, -1)
        for d in (F, V, dg, ds, drec):
            d.free()
    assert np.abs(rec).max() > 0
    assert np.array_equal(got_rec.view(np.uint32), rec.view(np.uint32))
    assert np.array_equal(gotP.view(np.uint8), P.view(np.uint8))
    assert np.array_equal(gotPP.view(np.uint8), PP.view(np.uint8))
