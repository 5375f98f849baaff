"""Parity at the benchmark's full grid sizes: one complete backward step of
HBC migration (source reconstruction with upper-border restore, receiver
tapers, fused receiver step + injection + imaging) on the 512x512x256+2*40
grid, four HBC forward steps and four modeling steps there with the FP32
defaults (tapers applied to each step's output), one stencil step on the
800x800x400+2*40 grid, and two fused HBC forward steps with FP16 and with
BF16 wavefields on that grid, bit-exact against the oracle on the same
random fields."""
import numpy as np
import pytest

from rtm3d_demo_amd import host

pytestmark = pytest.mark.gpu


def _bits_equal(a, b):
    return np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))


def _rand(shape, seed, lo=-1.0, hi=1.0):
    rng = np.random.default_rng(seed)
    out = np.empty(shape, np.float32)
    for x in range(shape[0]):  # plane by plane: bounded temporaries
        out[x] = rng.uniform(lo, hi, shape[1:]).astype(np.float32)
    return out


def test_backward_step_full_grid(gpu, oracle):
    nx, ny, nz, b, order = 512, 512, 256, 40, 8
    shape = (nx + 2 * b, ny + 2 * b, nz + 2 * b)
    print("full-size backward step", shape, flush=True)
    nt = 3
    Ps, PPs, Pr, PPr = (_rand(shape, s) for s in (1, 2, 3, 4))
    v2 = _rand(shape, 5, 0.0, 0.2)
    img = _rand(shape, 6)
    coefs = host.stencil_coefs(order, 12.5, 12.5, 12.5)
    g = host.taper_table(b, 0.9)
    upb = _rand((nt, shape[0], shape[1], 4), 7)
    geom = gpu.RcvGeom(0, 0, 2, 4, 4, 128, 128)
    rcv = _rand((128, 128, nt), 8)
    d = {k: gpu.DeviceArray.from_host(a) for k, a in
         dict(Ps=Ps, PPs=PPs, Pr=Pr, PPr=PPr, V=v2, img=img, upb=upb, rcv=rcv, g=g).items()}
    F, R = gpu.FieldPair.__new__(gpu.FieldPair), gpu.FieldPair.__new__(gpu.FieldPair)
    for fp, a, bb in ((F, d["Ps"], d["PPs"]), (R, d["Pr"], d["PPr"])):
        fp.a, fp.b = a, bb
        fp.p, fp.pp = type(a.ptr)(a.ptr.value), type(bb.ptr)(bb.ptr.value)
    shot = gpu.make_shot(*shape, b, order, nt, nt, (256 + b, 256 + b, 2 + b), geom)
    gpu.backward_steps(shot, True, d["V"], coefs, d["g"], F, R, d["upb"], d["rcv"], d["img"],
                       0, 1, True)
    # oracle: the same step (it = 0 <= nt-3)
    oracle.step(Ps, PPs, v2, coefs, order)
    oracle.upb(PPs, upb, b, order, 0, True)
    S = PPs
    oracle.taper(Pr, g, b, True)
    oracle.taper(PPr, g, b, True)
    oracle.step(Pr, PPr, v2, coefs, order)
    og = oracle.RcvGeom(0, 0, 2, 4, 4, 128, 128)
    oracle.receivers(PPr, rcv, b, og, nt, 0, True)
    oracle.image(img, S, PPr, b)
    assert _bits_equal(F.P.download(), S)
    assert _bits_equal(R.P.download(), PPr)
    assert _bits_equal(R.PP.download(), Pr)
    assert _bits_equal(d["img"].download(), img)
    for x in d.values():
        x.free()


SALT = (592, 592, 336)


def test_forward_steps_output_taper_full_grid(gpu, oracle):
    """Four HBC forward steps with the FP32 default, the upper taper applied
    to each step's output (the first step tapers P in place, the middle two
    run the output-taper kernel, the last the plain step), against the
    reference's sequence (taper, taper, step, source, save UPB)."""
    shape, b, order, nt = SALT, 40, 8, 4
    print("full-size forward steps", shape, flush=True)
    gpu.set_taper_fusion("f32", -1)
    P0, PP0 = _rand(shape, 31), _rand(shape, 32)
    v2 = _rand(shape, 33, 0.0, 0.2)
    g = host.taper_table(b, 0.9)
    coefs = host.stencil_coefs(order, 12.5, 12.5, 12.5)
    src = np.array([0.75, -0.5, 0.25, 1.0], np.float32)
    F = gpu.FieldPair(shape)
    F.a.upload(P0)
    F.b.upload(PP0)
    V, dg, ds = (gpu.DeviceArray.from_host(a) for a in (v2, g, src))
    dupb = gpu.DeviceArray((nt, shape[0], shape[1], 4)).zero()
    shot = gpu.make_shot(*shape, b, order, nt, nt, (296, 296, b + 2),
                         gpu.RcvGeom(0, 0, 0, 1, 1, 1, 1))
    gpu.forward_steps(shot, True, V, coefs, dg, ds, F, dupb, 0, nt)
    got_P, got_PP, got_upb = F.P.download(), F.PP.download(), dupb.download()
    for d in (F, V, dg, ds, dupb):
        d.free()
    P, PP = P0, PP0
    upb = np.zeros((nt, shape[0], shape[1], 4), np.float32)
    for it in range(nt):
        P, PP = PP, P
        oracle.taper(P, g, b, True)
        oracle.taper(PP, g, b, True)
        oracle.step(P, PP, v2, coefs, order)
        oracle.apply_source(PP, 296, 296, b + 2, src[it])
        oracle.upb(PP, upb, b, order, it, False)
    assert _bits_equal(got_PP, PP)
    assert _bits_equal(got_P, P)
    assert _bits_equal(got_upb, upb)


def test_model_steps_output_taper_full_grid(gpu, oracle):
    """Four modeling steps with the default, taper3D of both fields applied
    to each step's output (x-y frame and front-right edge by their own
    kernels), with the salt shot's 128 x 128 receivers, against the
    reference's sequence (taper3D, taper3D, step, source, save receivers)."""
    shape, b, order, nt = SALT, 40, 8, 4
    print("full-size modeling steps", shape, flush=True)
    P0, PP0 = _rand(shape, 41), _rand(shape, 42)
    v2 = _rand(shape, 43, 0.0, 0.2)
    g = host.taper_table(b, 0.9)
    coefs = host.stencil_coefs(order, 12.5, 12.5, 12.5)
    src = np.array([0.75, -0.5, 0.25, 1.0], np.float32)
    rgeom = (0, 0, 0, 4, 4, 128, 128)
    F = gpu.FieldPair(shape)
    F.a.upload(P0)
    F.b.upload(PP0)
    V, dg, ds = (gpu.DeviceArray.from_host(a) for a in (v2, g, src))
    drec = gpu.DeviceArray((128, 128, nt)).zero()
    shot = gpu.make_shot(*shape, b, order, nt, nt, (296, 296, b + 2), gpu.RcvGeom(*rgeom))
    gpu.model_steps(shot, V, coefs, dg, ds, F, drec, 0, nt)
    got_P, got_PP, got_rec = F.P.download(), F.PP.download(), drec.download()
    for d in (F, V, dg, ds, drec):
        d.free()
    P, PP = P0, PP0
    rec = np.zeros((128, 128, nt), np.float32)
    og = oracle.RcvGeom(*rgeom)
    for it in range(nt):
        P, PP = PP, P
        oracle.taper(P, g, b, False)
        oracle.taper(PP, g, b, False)
        oracle.step(P, PP, v2, coefs, order)
        oracle.apply_source(PP, 296, 296, b + 2, src[it])
        oracle.receivers(PP, rec, b, og, nt, it, False)
    assert _bits_equal(got_rec, rec)
    assert _bits_equal(got_PP, PP)
    assert _bits_equal(got_P, P)


def test_stencil_step_800_grid(gpu, oracle):
    shape = (880, 880, 480)
    print("stencil step", shape, flush=True)
    P, PP = _rand(shape, 11), _rand(shape, 12)
    V = _rand(shape, 13, 0.0, 0.2)
    coefs = host.stencil_coefs(8, 12.5, 12.5, 12.5)
    dP, dPP, dV = (gpu.DeviceArray.from_host(a) for a in (P, PP, V))
    del V
    gpu.step(dP, dPP, dV, coefs, 8)
    got = dPP.download()
    V = dV.download()
    oracle.step(P, PP, V, coefs, 8)
    assert _bits_equal(got, PP)
    for x in (dP, dPP, dV):
        x.free()


@pytest.mark.parametrize("fmt", ["f16", "bf16"])
def test_16bit_forward_steps_800_grid(gpu, oracle, fmt):
    """Two fused HBC forward steps with 16-bit wavefields on the
    800x800x400 + 2*40 grid (the mixed-precision configuration), bit-exact
    against the 16-bit oracle's operator sequence."""
    shape, b, order, nt = (880, 880, 480), 40, 8, 2
    print("16-bit forward steps", fmt, shape, flush=True)
    P0 = oracle.to_format(_rand(shape, 21), fmt)
    PP0 = oracle.to_format(_rand(shape, 22), fmt)
    v2 = _rand(shape, 23, 0.0, 0.2)
    g = host.taper_table(b, 0.9)
    coefs = host.stencil_coefs(order, 12.5, 12.5, 12.5)
    src = np.array([0.75, -0.5], np.float32)
    F = gpu.FieldPair(shape, np.uint16)
    F.a.upload(P0)
    F.b.upload(PP0)
    V, dg, ds = (gpu.DeviceArray.from_host(a) for a in (v2, g, src))
    dupb = gpu.DeviceArray((nt, shape[0], shape[1], 4), np.uint16).zero()
    shot = gpu.make_shot(*shape, b, order, nt, nt, (440, 440, b + 2),
                         gpu.RcvGeom(0, 0, 0, 1, 1, 1, 1), fmt)
    gpu.forward_steps(shot, True, V, coefs, dg, ds, F, dupb, 0, nt)
    got_P, got_PP, got_upb = F.P.download(), F.PP.download(), dupb.download()
    for d in (F, V, dg, ds, dupb):
        d.free()
    P, PP = P0, PP0
    upb = np.zeros((nt, shape[0], shape[1], 4), np.uint16)
    for it in range(nt):
        P, PP = PP, P
        oracle.taper(P, g, b, True, fmt)
        oracle.taper(PP, g, b, True, fmt)
        oracle.step(P, PP, v2, coefs, order, fmt)
        oracle.apply_source(PP, 440, 440, b + 2, src[it], fmt)
        oracle.upb(PP, upb, b, order, it, False, fmt)
    assert np.array_equal(got_PP, PP)
    assert np.array_equal(got_P, P)
    assert np.array_equal(got_upb, upb)
