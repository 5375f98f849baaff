"""Per-operator parity of the gfx950 kernels with the oracle, through the C
ABI: bit-exact FP32 on random fields, non-cubic grids whose sizes do not
divide the tile sizes, every supported stencil order."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(37, 45, 70), (12, 13, 11), (20, 91, 130), (70, 29, 33)]


def _rand(shape, seed, lo=-1.0, hi=1.0):
    rng = np.random.default_rng(seed)
    return rng.uniform(lo, hi, shape).astype(np.float32)


def _bits_equal(a, b):
    return np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))


@pytest.mark.parametrize("order", [2, 4, 6, 8])
@pytest.mark.parametrize("shape", SHAPES)
def test_step(gpu, oracle, order, shape):
    print("step", order, shape, flush=True)
    P, PP = _rand(shape, 1), _rand(shape, 2)
    V = _rand(shape, 3, 0.001, 0.2)
    coefs = oracle.stencil_coefs(order, 10.0, 12.0, 8.0)
    dP, dPP, dV = (gpu.DeviceArray.from_host(a) for a in (P, PP, V))
    gpu.step(dP, dPP, dV, coefs, order)
    want = PP.copy()
    assert oracle.step(P, want, V, coefs, order) == 0
    got = dPP.download()
    assert _bits_equal(got, want)
    assert _bits_equal(dP.download(), P)  # P is read-only
    for d in (dP, dPP, dV):
        d.free()


def test_step_zero_and_denormal_inputs(gpu, oracle):
    """Denormals are kept (no flush-to-zero), as on the reference's x86 build."""
    shape = (24, 26, 40)
    rng = np.random.default_rng(9)
    P = (rng.standard_normal(shape) * 1e-38).astype(np.float32)
    PP = (rng.standard_normal(shape) * 1e-39).astype(np.float32)
    V = np.full(shape, 0.1, np.float32)
    coefs = oracle.stencil_coefs(8, 1.0, 1.0, 1.0)
    dP, dPP, dV = (gpu.DeviceArray.from_host(a) for a in (P, PP, V))
    gpu.step(dP, dPP, dV, coefs, 8)
    want = PP.copy()
    oracle.step(P, want, V, coefs, 8)
    got = dPP.download()
    assert np.count_nonzero((np.abs(want) < 1.17e-38) & (want != 0)) > 100
    assert _bits_equal(got, want)


def test_step_nan_raises_dispersion(gpu, oracle):
    shape = (16, 16, 16)
    P = np.zeros(shape, np.float32)
    P[8, 8, 8] = np.nan
    dP, dPP, dV = (gpu.DeviceArray.from_host(a) for a in
                   (P, np.zeros(shape, np.float32), np.ones(shape, np.float32)))
    gpu.step(dP, dPP, dV, oracle.stencil_coefs(8, 1, 1, 1), 8)
    with pytest.raises(gpu.RtmError) as e:
        gpu.synchronize()
    assert e.value.code == gpu.ERR_DISPERSION and "Numerical Dispersion" in str(e.value)
    gpu.synchronize()  # flag cleared


def test_invalid_arguments(gpu, oracle):
    shape = (16, 16, 16)
    d = gpu.DeviceArray(shape).zero()
    with pytest.raises(gpu.RtmError) as e:
        gpu.step(d, d, d, np.zeros(33, np.float32), 10)
    assert e.value.code == gpu.ERR_UNSUPPORTED
    with pytest.raises(gpu.RtmError) as e:
        gpu.step(d, d, d, oracle.stencil_coefs(8, 1, 1, 1), 8)  # P == PP
    assert e.value.code == gpu.ERR_INVALID
    g = gpu.DeviceArray((8,)).zero()
    with pytest.raises(gpu.RtmError):
        gpu.taper(d, g, 8, True)  # blen leaves no interior
    d.free()
    g.free()


@pytest.mark.parametrize("upper", [False, True])
@pytest.mark.parametrize("shape,blen", [((37, 45, 70), 10), ((30, 22, 26), 4), ((64, 70, 90), 20),
                                        ((64, 70, 96), 16), ((30, 34, 40), 8)])
def test_taper(gpu, oracle, upper, shape, blen):
    P = _rand(shape, 4)
    g = oracle.taper_table(blen, np.float32(0.7))
    dP, dg = gpu.DeviceArray.from_host(P), gpu.DeviceArray.from_host(g)
    gpu.taper(dP, dg, blen, upper)
    want = P.copy()
    oracle.taper(want, g, blen, upper)
    assert _bits_equal(dP.download(), want)
    dP.free()
    dg.free()


def test_source_receivers_upb(gpu, oracle):
    shape, b = (37, 45, 70), 10
    P = _rand(shape, 5)
    dP = gpu.DeviceArray.from_host(P)
    gpu.apply_source(dP, 17, 21, 13, 0.625)
    want = P.copy()
    oracle.apply_source(want, 17, 21, 13, np.float32(0.625))
    assert _bits_equal(dP.download(), want)
    # receivers: save at lt=3 of a 7-long chunk, then restore into zeros
    geom = gpu.RcvGeom(1, 2, 1, 2, 3, 9, 8)
    ogeom = oracle.RcvGeom(1, 2, 1, 2, 3, 9, 8)
    rcv = _rand((9, 8, 7), 6)
    drcv = gpu.DeviceArray.from_host(rcv)
    gpu.receivers(dP, drcv, b, geom, 7, 3, False)
    wr = rcv.copy()
    oracle.receivers(want, wr, b, ogeom, 7, 3, False)
    assert _bits_equal(drcv.download(), wr)
    gpu.receivers(dP, drcv, b, geom, 7, 5, True)
    oracle.receivers(want, wr, b, ogeom, 7, 5, True)
    assert _bits_equal(dP.download(), want)
    # upper border save / restore with 64-bit offsets into a 3-step history
    upb = np.zeros((3, shape[0], shape[1], 4), np.float32)
    dupb = gpu.DeviceArray.from_host(upb)
    gpu.upper_border(dP, dupb, b, 8, 2, False)
    oracle.upb(want, upb, b, 8, 2, False)
    assert _bits_equal(dupb.download(), upb)
    Z = gpu.DeviceArray(shape).zero()
    gpu.upper_border(Z, dupb, b, 8, 2, True)
    wz = np.zeros(shape, np.float32)
    oracle.upb(wz, upb, b, 8, 2, True)
    assert _bits_equal(Z.download(), wz)
    for d in (dP, drcv, dupb, Z):
        d.free()


def test_image_stack_finish(gpu, oracle):
    shape, b = (37, 45, 70), 10
    img, PS, PR = _rand(shape, 7), _rand(shape, 8), _rand(shape, 9)
    d = [gpu.DeviceArray.from_host(a) for a in (img, PS, PR)]
    gpu.image_condition(d[0], d[1], d[2], b)
    want = img.copy()
    oracle.image(want, PS, PR, b)
    assert _bits_equal(d[0].download(), want)
    gpu.stack(d[0], d[1])
    want = want + PS
    assert _bits_equal(d[0].download(), want)
    fc = oracle.stencil_coefs(2, 10.0, 12.0, 8.0)
    out = gpu.DeviceArray((shape[0] - 2 * b, shape[1] - 2 * b, shape[2] - 2 * b))
    gpu.image_finish(d[0], out, b, fc)
    assert _bits_equal(out.download(), oracle.image_finish(want, b, 10.0, 12.0, 8.0))
    for x in d + [out]:
        x.free()


@pytest.mark.parametrize("nwstep, iw0", [(5, 0), (3, 2), (4, 3)])
def test_freq_image_matches_oracle(gpu, oracle, nwstep, iw0):
    """Frequency contributions + imaging (rtmUpdateFreqContributions and
    rtmFreqDomainImageCondition) for a group of frequencies, bit-exact
    against the oracle on non-zero accumulators; the last group is cut at
    nw."""
    rng = np.random.default_rng(40 + iw0)
    shape, b, nt, st = (30, 34, 28), 6, 40, 17
    nw = 5
    inner = tuple(n - 2 * b for n in shape)
    PS, PR, img = (rng.uniform(-1, 1, shape).astype(np.float32) for _ in range(3))
    acc = [rng.uniform(-2, 2, (nwstep,) + inner).astype(np.float32) for _ in range(4)]
    w2, kre, kim = oracle.freq_tables(nt, np.float32(0.004), np.float32(2.0), np.float32(40.0))
    assert len(w2) >= nw  # the kernel reads kre / kim rows iw0 .. nw-1
    w2, kre, kim = (np.ascontiguousarray(a[:nw]) for a in (w2, kre, kim))
    d = [gpu.DeviceArray.from_host(a) for a in [img, PS, PR] + acc + [kre, kim, w2]]
    gpu.freq_image(d[0], d[1], d[2], d[3:7], d[7], d[8], d[9], b, nwstep, iw0, nw, nt, st)
    oracle.freq_image(img, PS, PR, acc, kre, kim, w2, b, nwstep, iw0, nt, st)
    assert np.array_equal(d[0].download().view(np.uint32), img.view(np.uint32))
    for k in range(4):
        assert np.array_equal(d[3 + k].download().view(np.uint32), acc[k].view(np.uint32))
    for x in d:
        x.free()
