"""Device-resident time loops (modeling, HBC / RBC migration, fused and
unfused backward step) against the oracle and against the reference
program's fixtures, bit-exact; plus time-reversibility of the HBC
reconstruction on the device."""
import numpy as np
import pytest

from rtm3d_demo_amd import host

pytestmark = pytest.mark.gpu


def _bits_equal(a, b):
    return np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))


def _geom(gpu, p):
    return gpu.RcvGeom(p["receiver_start_x"], p["receiver_start_y"], p["receiver_depth_z"],
                       p["receiver_distance_x"], p["receiver_distance_y"],
                       p["receiver_count_x"], p["receiver_count_y"])


def _shots(p):
    return [(ix * p["source_distance_x"] + p["source_start_x"],
             iy * p["source_distance_y"] + p["source_start_y"], p["source_depth_z"])
            for ix in range(p["source_count_x"]) for iy in range(p["source_count_y"])]


class Setup:
    """Product host set-up of one case, uploaded to the device; fmt is the
    wavefield storage format ('f32', 'f16', 'bf16')."""

    def __init__(self, gpu, p, ve, fmt="f32"):
        self.gpu, self.p, self.fmt = gpu, p, fmt
        self.wdtype = gpu.WAVE_FORMATS[fmt][1]
        self.blen, self.order = p["blen"], p["stencil_order"]
        self.nt, self.ntstep = p["nt"], p["ntstep"]
        self.shape = ve.shape
        self.coefs = host.stencil_coefs(self.order, p["dx"], p["dy"], p["dz"])
        self.V = gpu.DeviceArray.from_host(host.v2dt2(ve, p["dt"]))
        self.src = gpu.DeviceArray.from_host(host.ricker(self.nt, p["dt"], p["fpeak"]))
        self.g = gpu.DeviceArray.from_host(host.taper_table(self.blen, p["taper_factor"]))
        self.geom = _geom(gpu, p)

    def shot(self, s):
        b = self.blen
        return self.gpu.make_shot(*self.shape, b, self.order, self.nt, self.ntstep,
                                  (s[0] + b, s[1] + b, s[2] + b), self.geom, self.fmt)

    def free(self):
        for d in (self.V, self.src, self.g):
            d.free()


def gpu_records(gpu, p, vel, fmt="f32"):
    S = Setup(gpu, p, host.extend_velocity(vel, p["blen"], "abc"), fmt)
    out = []
    for s in _shots(p):
        F = gpu.FieldPair(S.shape, S.wdtype)
        rcv = gpu.DeviceArray((p["receiver_count_x"], p["receiver_count_y"], S.ntstep)).zero()
        chunks = []
        for itstep in range(0, S.nt, S.ntstep):
            gpu.model_steps(S.shot(s), S.V, S.coefs, S.g, S.src, F, rcv, itstep, S.ntstep)
            chunks.append(rcv.download())
        out.append(np.stack(chunks))
        F.free()
        rcv.free()
    S.free()
    return np.stack(out)


def gpu_shot_image(gpu, S, s, records, hbc, fused, spill=False):
    """spill: the HBC history of one receiver chunk on the device (hbc mode
    2), each chunk's layers kept on the host between the passes, as the
    reference spills them to files."""
    b = S.blen
    F, R = gpu.FieldPair(S.shape, S.wdtype), gpu.FieldPair(S.shape, S.wdtype)
    mode = (2 if spill else 1) if hbc else 0
    upb = (gpu.DeviceArray((S.ntstep if spill else S.nt, S.shape[0], S.shape[1], S.order // 2),
                           S.wdtype) if hbc else None)
    spilled = {}
    for itstep in range(0, S.nt, S.ntstep):
        if spill:
            upb.upload(np.full(upb.shape, 0x7FC0, np.uint16).view(S.wdtype)
                       if S.wdtype == np.uint16 else np.full(upb.shape, np.nan, np.float32))
        gpu.forward_steps(S.shot(s), mode, S.V, S.coefs, S.g, S.src, F, upb, itstep, S.ntstep)
        if spill:
            spilled[itstep] = upb.download()
    img = gpu.DeviceArray(S.shape).zero()
    rcv = gpu.DeviceArray(records.shape[1:])
    for k, itstep in reversed(list(enumerate(range(0, S.nt, S.ntstep)))):
        rcv.upload(records[k])
        if spill:
            upb.upload(spilled[itstep])
        gpu.backward_steps(S.shot(s), mode, S.V, S.coefs, S.g, F, R, upb, rcv, img,
                           itstep + S.ntstep - 1, S.ntstep, fused)
    out = img.download()
    for d in (F, R, img, rcv) + ((upb,) if upb else ()):
        d.free()
    return out


def gpu_migration(gpu, p, vpe, records, fused=True, fmt="f32", spill=False):
    hbc = p["boundary_condition"] != "rbc"
    S = Setup(gpu, p, vpe, fmt)
    total = gpu.DeviceArray(S.shape).zero()
    simg = gpu.DeviceArray(S.shape)
    for k, s in enumerate(_shots(p)):
        simg.upload(gpu_shot_image(gpu, S, s, records[k], hbc, fused, spill))
        gpu.stack(total, simg)
    b = S.blen
    out = gpu.DeviceArray(tuple(n - 2 * b for n in S.shape))
    gpu.image_finish(total, out, b, host.stencil_coefs(2, p["dx"], p["dy"], p["dz"]))
    img = out.download()
    for d in (total, simg, out):
        d.free()
    S.free()
    return img


@pytest.mark.parametrize("name", ["small_mod", "o4_mod"])
def test_modeling_matches_reference(gpu, cases, golden, name):
    print("modeling", name, flush=True)
    p, fx = cases[name], golden(name)
    rec = gpu_records(gpu, p, fx["vel"])
    assert _bits_equal(rec, fx["records"])


@pytest.mark.parametrize("fused", [True, False, 2])
@pytest.mark.parametrize("bc", ["hbc", "rbc"])
def test_migration_matches_reference(gpu, cases, golden, bc, fused):
    print("migration", bc, "fused" if fused else "unfused", flush=True)
    p = cases["small_" + bc]
    img = gpu_migration(gpu, p, golden("small_" + bc)["vpe"], golden("small_mod")["records"], fused)
    assert _bits_equal(img, golden("small_" + bc)["image"])


@pytest.mark.parametrize("fused", [True, False, 2])
def test_migration_one_chunk_history_matches_reference(gpu, cases, golden, fused):
    """HBC history of one receiver chunk on the device (hbc mode 2, the
    reference's per-chunk upper-border grid; the buffer is filled with NaN
    before each forward chunk, so a layer read from the wrong slot shows)."""
    print("migration one-chunk history", fused, flush=True)
    p = cases["small_hbc"]
    assert p["nt"] // p["ntstep"] >= 2
    img = gpu_migration(gpu, p, golden("small_hbc")["vpe"], golden("small_mod")["records"],
                        fused, spill=True)
    assert _bits_equal(img, golden("small_hbc")["image"])


@pytest.mark.parametrize("fmt", ["f16", "bf16"])
def test_one_chunk_history_16bit(gpu, cases, golden, fmt):
    """16-bit wavefields and history: mode 2 gives the whole-history image."""
    print("one-chunk history", fmt, flush=True)
    p = cases["small_hbc"]
    vpe, rec = golden("small_hbc")["vpe"], golden("small_mod")["records"]
    a = gpu_migration(gpu, p, vpe, rec, True, fmt)
    b = gpu_migration(gpu, p, vpe, rec, True, fmt, spill=True)
    assert _bits_equal(a, b)


def test_one_chunk_history_rejects_bad_calls(gpu, cases, golden):
    p = cases["small_hbc"]
    S = Setup(gpu, p, golden("small_hbc")["vpe"])
    s = _shots(p)[0]
    F = gpu.FieldPair(S.shape)
    upb = gpu.DeviceArray((S.ntstep, S.shape[0], S.shape[1], S.order // 2))
    with pytest.raises(Exception, match="crosses a receiver chunk"):
        gpu.forward_steps(S.shot(s), 2, S.V, S.coefs, S.g, S.src, F, upb, S.ntstep - 2, 4)
    with pytest.raises(Exception, match="hbc must be"):
        gpu.forward_steps(S.shot(s), 3, S.V, S.coefs, S.g, S.src, F, upb, 0, 2)
    for d in (F, upb):
        d.free()
    S.free()


def test_migration_matches_oracle_other_geometry(gpu, oracle, cases, golden):
    """Receivers deeper than the source, one-step chunks at the end, order 6."""
    print("migration order 6", flush=True)
    p = dict(cases["small_hbc"])
    p.update({"stencil_order": 6, "nt": 90, "ntstep": 30, "receiver_depth_z": 5,
              "receiver_start_x": 0, "receiver_distance_x": 1, "receiver_count_x": 36,
              "source_count_x": 1, "source_start_x": 3, "source_start_y": 40})
    vel = golden("small_mod")["vel"]
    mp = dict(p, modeling=True)
    rec = gpu_records(gpu, mp, vel)
    assert _bits_equal(rec, oracle.pipeline_modeling(mp, vel))
    vpe = host.extend_velocity(vel, p["blen"], "hbc")
    img = gpu_migration(gpu, p, vpe, rec)
    assert _bits_equal(img, oracle.pipeline_migration(p, vpe, rec))


def test_hbc_reconstruction_is_time_reversible(gpu):
    """No taper (factor 1), source silent after step K: the backward source
    field at time K equals the forward field at K below z = blen-4, to FP32
    rounding accumulated over the reversal."""
    shape, b, order, nt, K = (48, 50, 64), 10, 8, 160, 60
    rng = np.random.default_rng(0)
    v2 = rng.uniform(0.02, 0.12, shape).astype(np.float32)
    V = gpu.DeviceArray.from_host(v2)
    coefs = host.stencil_coefs(order, 1.0, 1.0, 1.0)
    g = gpu.DeviceArray.from_host(np.ones(b, np.float32))
    s = host.ricker(nt, 0.004, 20.0)
    s[K:] = 0
    src = gpu.DeviceArray.from_host(s)
    geom = gpu.RcvGeom(0, 0, 0, 1, 1, 1, 1)
    shot = gpu.make_shot(*shape, b, order, nt, nt, (24, 25, b + 3), geom)
    F = gpu.FieldPair(shape)
    upb = gpu.DeviceArray((nt, shape[0], shape[1], 4))
    gpu.forward_steps(shot, True, V, coefs, g, src, F, upb, 0, K + 1)
    pK = F.PP.download()
    gpu.forward_steps(shot, True, V, coefs, g, src, F, upb, K + 1, nt - K - 1)
    R = gpu.FieldPair(shape)
    rcv = gpu.DeviceArray((1, 1, nt)).zero()
    img = gpu.DeviceArray(shape).zero()
    gpu.backward_steps(shot, True, V, coefs, g, F, R, upb, rcv, img, nt - 1, nt - K, True)
    rec = F.P.download()
    err = np.abs(rec - pK)[:, :, b - 4:].max() / np.abs(pK).max()
    assert np.abs(pK).max() > 1e-3
    assert err < 1e-4, err
    assert np.count_nonzero(img.download()) == 0  # zero receiver field images nothing
    for d in (V, g, src, F, upb, R, rcv, img):
        d.free()


def _pair(gpu, A, B):
    F = gpu.FieldPair(A.shape)
    F.a.upload(A)
    F.b.upload(B)
    return F


@pytest.fixture(params=[((0, 0, 0), 0), ((1, 4, 4), 1), ((1, 4, 4), 5), ((1, 2, 6), 0),
                        ((1, 2, 7), 3), ((2, 4, 3), 0), ((2, 4, 3), 5), ((64, 1, 6), 0),
                        ((64, 1, 8), 3), ((128, 1, 4), 0), ((128, 1, 4), 3), ((128, 1, 6), 0),
                        ((128, 1, 8), 3), ((64, 4, 4), 0), ((64, 4, 4), 5)],
                ids=lambda p: "t%d.%d.%d-x%d" % (*p[0], p[1]))
def xchunk(request, gpu):
    """Stencil shape and planes marched per block: the built-in default, each
    full-row shape and the z-tiled shape, automatic or fixed short marches
    (multi-plane marching through the register ring on small grids)."""
    tiling, xc = request.param
    for cls in range(3):
        gpu.set_tiling(cls, *tiling, xc)
    yield request.param
    for cls in range(3):
        gpu.set_tiling(cls, 0)


# (40, 44, 90): two waves per z row; (32, 36, 340): six waves per z row;
# (30, 32, 420): seven, beyond every full-row and pair shape, so those
# requests take the fallback (64x1x8 waves); (30, 34, 91): odd nz, so the
# two-z-per-lane shapes run one z per lane.
GRIDS = [(40, 44, 90), (32, 36, 340), (30, 32, 420), (30, 34, 91)]


@pytest.fixture(params=[1, 0, 4], ids=["tfused", "tsep", "tout"])
def fusion32(request, gpu):
    """Upper tapers fused into the FP32 stencil (the default), applied by a
    separate in-place pass, or applied to each step's output."""
    gpu.set_taper_fusion("f32", request.param)
    yield request.param
    gpu.set_taper_fusion("f32", -1)


@pytest.mark.parametrize("b", [12, 11])
@pytest.mark.parametrize("shape", GRIDS)
@pytest.mark.parametrize("splits", [(5,), (2, 1, 2)])
def test_fused_taper_forward_matches_operator_sequence(gpu, oracle, splits, xchunk, shape, b,
                                                       fusion32):
    """HBC forward steps with the upper taper fused into the stencil (or in a
    separate pass) equal the reference's operator sequence (taper, taper,
    step, source, save UPB) on non-zero fields, whether run in one call or
    split into calls.  An odd border puts the taper edge inside a z pair of
    the two-z kernels."""
    order, nt = 8, 5
    rng = np.random.default_rng(21)
    P0 = rng.uniform(-1, 1, shape).astype(np.float32)
    PP0 = rng.uniform(-1, 1, shape).astype(np.float32)
    v2 = rng.uniform(0.0, 0.15, shape).astype(np.float32)
    g = oracle.taper_table(b, np.float32(0.3))
    coefs = oracle.stencil_coefs(order, 10.0, 11.0, 9.0)
    src = rng.uniform(-1, 1, nt).astype(np.float32)
    # oracle: reference operator order
    P, PP = P0.copy(), PP0.copy()
    upb = np.zeros((nt, shape[0], shape[1], 4), np.float32)
    for it in range(nt):
        P, PP = PP, P
        oracle.taper(P, g, b, True)
        oracle.taper(PP, g, b, True)
        oracle.step(P, PP, v2, coefs, order)
        oracle.apply_source(PP, 20, 22, b + 2, src[it])
        oracle.upb(PP, upb, b, order, it, False)
    F = _pair(gpu, P0, PP0)
    V, dg, ds = (gpu.DeviceArray.from_host(a) for a in (v2, g, src))
    dupb = gpu.DeviceArray(upb.shape).zero()
    shot = gpu.make_shot(*shape, b, order, nt, nt, (20, 22, b + 2), gpu.RcvGeom(0, 0, 0, 1, 1, 1, 1))
    it0 = 0
    for n in splits:
        gpu.forward_steps(shot, True, V, coefs, dg, ds, F, dupb, it0, n)
        it0 += n
    assert _bits_equal(F.P.download(), P)
    assert _bits_equal(F.PP.download(), PP)
    assert _bits_equal(dupb.download(), upb)
    for d in (F, V, dg, ds, dupb):
        d.free()


@pytest.mark.parametrize("b", [12, 11])
@pytest.mark.parametrize("mode, fusion", [(1, 1), (1, 0), (1, 4), (2, 1)],
                         ids=["m1-tfused", "m1-tsep", "m1-tout", "m2"])
@pytest.mark.parametrize("shape", GRIDS)
@pytest.mark.parametrize("splits", [(3,), (1, 2)])
def test_fused_backward_matches_operator_sequence(gpu, oracle, splits, xchunk, shape, mode,
                                                  fusion, b):
    """Fused receiver step (tapers, stencil, receiver injection, imaging) --
    with the tapers fused or in a separate pass, and with mode 2 the source
    field's backward step in the same pass -- against the reference's
    operator sequence on non-zero fields; odd and even borders."""
    gpu.set_taper_fusion("f32", fusion)
    order, nt = 8, 8
    rng = np.random.default_rng(22)
    Ps, PPs, Pr, PPr, img = (rng.uniform(-1, 1, shape).astype(np.float32) for _ in range(5))
    v2 = rng.uniform(0.0, 0.15, shape).astype(np.float32)
    g = oracle.taper_table(b, np.float32(0.3))
    coefs = oracle.stencil_coefs(order, 10.0, 11.0, 9.0)
    upb = rng.uniform(-1, 1, (nt, shape[0], shape[1], 4)).astype(np.float32)
    rcv = rng.uniform(-1, 1, (7, 6, nt)).astype(np.float32)
    og = oracle.RcvGeom(1, 2, 3, 3, 4, 7, 6)
    geom = gpu.RcvGeom(1, 2, 3, 3, 4, 7, 6)
    d = {k: gpu.DeviceArray.from_host(a) for k, a in
         dict(V=v2, g=g, upb=upb, rcv=rcv, img=img).items()}
    S, R = _pair(gpu, Ps, PPs), _pair(gpu, Pr, PPr)
    shot = gpu.make_shot(*shape, b, order, nt, nt, (20, 22, b + 2), geom)
    hi = nt - 3
    for n in splits:
        gpu.backward_steps(shot, True, d["V"], coefs, d["g"], S, R, d["upb"], d["rcv"],
                           d["img"], hi, n, mode)
        hi -= n
    for it in range(nt - 3, nt - 3 - sum(splits), -1):
        oracle.step(Ps, PPs, v2, coefs, order)
        oracle.upb(PPs, upb, b, order, it, True)
        Ps, PPs = PPs, Ps
        oracle.taper(Pr, g, b, True)
        oracle.taper(PPr, g, b, True)
        oracle.step(Pr, PPr, v2, coefs, order)
        oracle.receivers(PPr, rcv, b, og, nt, it, True)
        oracle.image(img, Ps, PPr, b)
        Pr, PPr = PPr, Pr
    gpu.set_taper_fusion("f32", -1)
    assert _bits_equal(S.P.download(), Ps) and _bits_equal(S.PP.download(), PPs)
    assert _bits_equal(R.P.download(), Pr) and _bits_equal(R.PP.download(), PPr)
    assert _bits_equal(d["img"].download(), img)
    for x in list(d.values()) + [S, R]:
        x.free()


@pytest.mark.parametrize("tiling", [(64, 4, 4), (64, 4, 2), (64, 8, 2), (32, 8, 4), (32, 8, 2),
                                    (1, 4, 4), (1, 2, 6), (1, 2, 7), (2, 4, 3), (64, 1, 6),
                                    (64, 1, 8), (128, 1, 4), (128, 1, 6), (128, 1, 8)])
def test_tile_shapes_bit_exact(gpu, oracle, cases, golden, tiling):
    """Every compiled tile shape, for all three stencil classes, reproduces
    the reference HBC image bit for bit."""
    for cls in range(3):
        gpu.set_tiling(cls, *tiling, 0)
    try:
        p = cases["small_hbc"]
        img = gpu_migration(gpu, p, golden("small_hbc")["vpe"], golden("small_mod")["records"])
        assert _bits_equal(img, golden("small_hbc")["image"])
    finally:
        for cls in range(3):
            gpu.set_tiling(cls, 0)
