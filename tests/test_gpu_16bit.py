"""16-bit wavefield storage (RTMHIP_WAVE_F16 / _BF16) on the device.

Bit for bit against the 16-bit oracle (the reference's operator sequence on
16-bit grids, float arithmetic, every wavefield write rounded): modeling
traces, HBC/RBC images (fused and unfused backward step), and the fused
forward / backward steps on non-zero fields for each stencil shape.  Then
the accuracy of each format against the FP64 oracle, with the measured
relative L2 errors written into the bounds.
"""
import numpy as np
import pytest

from rtm3d_demo_amd import host
from test_gpu_loops import gpu_migration, gpu_records

pytestmark = pytest.mark.gpu

FMT16 = ("f16", "bf16")


def _same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))


@pytest.mark.parametrize("fmt", FMT16)
def test_modeling_16bit_matches_oracle(gpu, oracle, cases, golden, fmt):
    print("modeling", fmt, flush=True)
    p = cases["small_mod"]
    vel = golden("small_mod")["vel"]
    rec = gpu_records(gpu, p, vel, fmt)
    assert _same(rec, oracle.pipeline_modeling(p, vel, np.float32, fmt))


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("bc", ["hbc", "rbc"])
@pytest.mark.parametrize("fmt", FMT16)
def test_migration_16bit_matches_oracle(gpu, oracle, cases, golden, fmt, bc, fused, fusion16):
    print("migration", fmt, bc, fused, flush=True)
    p = cases["small_" + bc]
    vpe = golden("small_" + bc)["vpe"]
    records = golden("small_mod")["records"]
    img = gpu_migration(gpu, p, vpe, records, fused, fmt)
    assert _same(img, oracle.pipeline_migration(p, vpe, records, np.float32, fmt))


@pytest.fixture(params=[(0, 0, 0), (1, 4, 4), (1, 2, 6), (64, 1, 6), (128, 1, 4), (128, 1, 8),
                        (64, 4, 4)],
                ids=lambda t: "t%d.%d.%d" % t)
def shape16(request, gpu):
    """Default shapes, full rows requested (run where they cover nz), waves
    requested, tiles requested (16-bit runs the default for tiles)."""
    for cls in range(3):
        gpu.set_tiling(cls, *request.param, 0)
    yield request.param
    for cls in range(3):
        gpu.set_tiling(cls, 0)


@pytest.fixture(params=[-1, 1, 4], ids=["tdefault", "tfused", "tout"])
def fusion16(request, gpu):
    """The 16-bit default (upper tapers in a separate in-place pass), the
    tapers fused into the stencil, or applied to each step's output."""
    for fmt in FMT16:
        gpu.set_taper_fusion(fmt, request.param)
    yield request.param
    for fmt in FMT16:
        gpu.set_taper_fusion(fmt, -1)


# two and six waves per z row (full-row kernel), seven (z-tiled kernel);
# odd nz: one z per lane for every wave shape
GRIDS16 = [(40, 44, 90), (32, 36, 340), (30, 32, 420), (30, 34, 91)]


@pytest.mark.parametrize("b", [12, 11])
@pytest.mark.parametrize("grid", GRIDS16)
@pytest.mark.parametrize("fmt", FMT16)
def test_fused_forward_16bit_matches_operator_sequence(gpu, oracle, fmt, grid, shape16, fusion16,
                                                       b):
    order, nt = 8, 4
    rng = np.random.default_rng(31)
    P0 = oracle.to_format(rng.uniform(-1, 1, grid).astype(np.float32), fmt)
    PP0 = oracle.to_format(rng.uniform(-1, 1, grid).astype(np.float32), fmt)
    v2 = rng.uniform(0.0, 0.15, grid).astype(np.float32)
    g = oracle.taper_table(b, np.float32(0.3))
    coefs = oracle.stencil_coefs(order, 10.0, 11.0, 9.0)
    src = rng.uniform(-1, 1, nt).astype(np.float32)
    P, PP = P0.copy(), PP0.copy()
    upb = np.zeros((nt, grid[0], grid[1], 4), np.uint16)
    for it in range(nt):
        P, PP = PP, P
        oracle.taper(P, g, b, True, fmt)
        oracle.taper(PP, g, b, True, fmt)
        oracle.step(P, PP, v2, coefs, order, fmt)
        oracle.apply_source(PP, 20, 22, b + 2, src[it], fmt)
        oracle.upb(PP, upb, b, order, it, False, fmt)
    F = gpu.FieldPair(grid, np.uint16)
    F.a.upload(P0)
    F.b.upload(PP0)
    V, dg, ds = (gpu.DeviceArray.from_host(a) for a in (v2, g, src))
    dupb = gpu.DeviceArray(upb.shape, np.uint16).zero()
    shot = gpu.make_shot(*grid, b, order, nt, nt, (20, 22, b + 2),
                         gpu.RcvGeom(0, 0, 0, 1, 1, 1, 1), fmt)
    gpu.forward_steps(shot, True, V, coefs, dg, ds, F, dupb, 0, 2)
    gpu.forward_steps(shot, True, V, coefs, dg, ds, F, dupb, 2, 2)
    assert np.array_equal(F.P.download(), P)
    assert np.array_equal(F.PP.download(), PP)
    assert np.array_equal(dupb.download(), upb)
    for d in (F, V, dg, ds, dupb):
        d.free()


@pytest.mark.parametrize("b", [12, 11])
@pytest.mark.parametrize("grid", GRIDS16)
@pytest.mark.parametrize("fmt", FMT16)
def test_fused_backward_16bit_matches_operator_sequence(gpu, oracle, fmt, grid, shape16, fusion16,
                                                        b):
    order, nt = 8, 8
    rng = np.random.default_rng(32)
    q = lambda: oracle.to_format(rng.uniform(-1, 1, grid).astype(np.float32), fmt)
    Ps, PPs, Pr, PPr = q(), q(), q(), q()
    img = rng.uniform(-1, 1, grid).astype(np.float32)
    v2 = rng.uniform(0.0, 0.15, grid).astype(np.float32)
    g = oracle.taper_table(b, np.float32(0.3))
    coefs = oracle.stencil_coefs(order, 10.0, 11.0, 9.0)
    upb = oracle.to_format(rng.uniform(-1, 1, (nt, grid[0], grid[1], 4)).astype(np.float32), fmt)
    rcv = rng.uniform(-1, 1, (7, 6, nt)).astype(np.float32)
    og = oracle.RcvGeom(1, 2, 3, 3, 4, 7, 6)
    d = {k: gpu.DeviceArray.from_host(a) for k, a in dict(V=v2, g=g, rcv=rcv, img=img).items()}
    d["upb"] = gpu.DeviceArray.from_host(upb, np.uint16)
    S, R = gpu.FieldPair(grid, np.uint16), gpu.FieldPair(grid, np.uint16)
    S.a.upload(Ps)
    S.b.upload(PPs)
    R.a.upload(Pr)
    R.b.upload(PPr)
    shot = gpu.make_shot(*grid, b, order, nt, nt, (20, 22, b + 2),
                         gpu.RcvGeom(1, 2, 3, 3, 4, 7, 6), fmt)
    hi = nt - 3
    gpu.backward_steps(shot, True, d["V"], coefs, d["g"], S, R, d["upb"], d["rcv"], d["img"], hi,
                       1, True)
    gpu.backward_steps(shot, True, d["V"], coefs, d["g"], S, R, d["upb"], d["rcv"], d["img"],
                       hi - 1, 2, True)
    for it in range(hi, hi - 3, -1):
        oracle.step(Ps, PPs, v2, coefs, order, fmt)
        oracle.upb(PPs, upb, b, order, it, True, fmt)
        Ps, PPs = PPs, Ps
        oracle.taper(Pr, g, b, True, fmt)
        oracle.taper(PPr, g, b, True, fmt)
        oracle.step(Pr, PPr, v2, coefs, order, fmt)
        oracle.receivers(PPr, rcv, b, og, nt, it, True, fmt)
        oracle.image(img, Ps, PPr, b, fmt)
        Pr, PPr = PPr, Pr
    assert np.array_equal(S.P.download(), Ps) and np.array_equal(S.PP.download(), PPs)
    assert np.array_equal(R.P.download(), Pr) and np.array_equal(R.PP.download(), PPr)
    assert _same(d["img"].download(), img)
    for x in list(d.values()) + [S, R]:
        x.free()


@pytest.mark.parametrize("fmt, bound", [("f32", 1e-5), ("f16", 1e-2), ("bf16", 1e-1)])
def test_image_error_vs_fp64(gpu, oracle, cases, golden, fmt, bound):
    """Final HBC image of the small case (two shots, 300 steps) from the
    device with each storage format, against the FP64 oracle image on the
    same records.  The 16-bit oracle gives 1.7e-6 (f32), 3.2e-3 (f16) and
    5.2e-2 (bf16): f16 meets the 1e-2 image target, bf16 (8-bit
    significand) does not, and its bound only guards against regressions."""
    p = cases["small_hbc"]
    vpe = golden("small_hbc")["vpe"]
    records = golden("small_mod")["records"]
    ref = oracle.pipeline_migration(p, vpe, records.astype(np.float64), np.float64)
    img = gpu_migration(gpu, p, vpe, records, True, fmt)
    err = float(np.linalg.norm(img - ref) / np.linalg.norm(ref))
    print("image rel-L2 vs FP64 (%s): %.3e" % (fmt, err), flush=True)
    assert err < bound


@pytest.mark.parametrize("fmt", ["f32", "f16", "bf16"])
def test_taper_vector_forms_match_oracle(gpu, oracle, fmt):
    """Border widths and nz that are multiples of the 16-byte group (4 FP32
    or 8 16-bit cells) take the grouped taper kernels: the all-border taper
    of the modeling loop and the separate upper-taper pass of the HBC loops,
    on non-zero fields, against the operator sequence."""
    grid, b, order, nt = (40, 44, 96), 16, 8, 4
    rng = np.random.default_rng(33)
    wdt = np.float32 if fmt == "f32" else np.uint16
    q = (lambda a: a) if fmt == "f32" else (lambda a: oracle.to_format(a, fmt))
    P0 = q(rng.uniform(-1, 1, grid).astype(np.float32))
    PP0 = q(rng.uniform(-1, 1, grid).astype(np.float32))
    v2 = rng.uniform(0.0, 0.15, grid).astype(np.float32)
    g = oracle.taper_table(b, np.float32(0.3))
    coefs = oracle.stencil_coefs(order, 10.0, 11.0, 9.0)
    src = rng.uniform(-1, 1, nt).astype(np.float32)
    ofmt = None if fmt == "f32" else fmt
    V, dg, ds = (gpu.DeviceArray.from_host(a) for a in (v2, g, src))
    shot = gpu.make_shot(*grid, b, order, nt, nt, (20, 22, b + 2),
                         gpu.RcvGeom(0, 0, 0, 1, 1, 1, 1), fmt)
    # modeling: taper all borders of both fields, step, source
    P, PP = P0.copy(), PP0.copy()
    rec = np.zeros((1, 1, nt), np.float32)
    for it in range(nt):
        P, PP = PP, P
        oracle.taper(P, g, b, False, ofmt)
        oracle.taper(PP, g, b, False, ofmt)
        oracle.step(P, PP, v2, coefs, order, ofmt)
        oracle.apply_source(PP, 20, 22, b + 2, src[it], ofmt)
    F = gpu.FieldPair(grid, wdt)
    F.a.upload(P0)
    F.b.upload(PP0)
    drec = gpu.DeviceArray((1, 1, nt)).zero()
    gpu.model_steps(shot, V, coefs, dg, ds, F, drec, 0, nt)
    assert np.array_equal(F.P.download(), P) and np.array_equal(F.PP.download(), PP)
    # HBC forward with the separate upper-taper pass
    gpu.set_taper_fusion(fmt, 0)
    try:
        P, PP = P0.copy(), PP0.copy()
        upb = np.zeros((nt, grid[0], grid[1], 4), wdt)
        for it in range(nt):
            P, PP = PP, P
            oracle.taper(P, g, b, True, ofmt)
            oracle.taper(PP, g, b, True, ofmt)
            oracle.step(P, PP, v2, coefs, order, ofmt)
            oracle.apply_source(PP, 20, 22, b + 2, src[it], ofmt)
            oracle.upb(PP, upb, b, order, it, False, ofmt)
        F.a.upload(P0)
        F.b.upload(PP0)
        F.p.value, F.pp.value = F.a.ptr.value, F.b.ptr.value
        dupb = gpu.DeviceArray(upb.shape, wdt).zero()
        gpu.forward_steps(shot, True, V, coefs, dg, ds, F, dupb, 0, nt)
        assert np.array_equal(F.P.download(), P) and np.array_equal(F.PP.download(), PP)
        assert np.array_equal(dupb.download(), upb)
        dupb.free()
    finally:
        gpu.set_taper_fusion(fmt, -1)
        for d in (F, V, dg, ds, drec):
            d.free()
