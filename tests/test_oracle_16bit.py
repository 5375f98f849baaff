"""Pins the 16-bit-storage oracle (orc_*_f16, orc_*_bf16).

* The format conversions against numpy's float16 casts (every one of the
  65536 half values widened; random and edge floats narrowed) and torch's
  bfloat16 casts.
* Every operator that rounds once per written cell equals "round(the float
  operator on the widened field)", cell for cell: the 16-bit build is the
  float build (pinned to the reference program's own outputs) with the
  storage rounding at exactly the wavefield writes.
* The front-right taper quirk reads a face-tapered cell that the 16-bit
  grid has already rounded; checked by hand on a tiny grid.
* Whole modeling runs: trace error against the FP64 oracle, per format.
"""
import numpy as np
import pytest

FMT16 = ("f16", "bf16")


def _bits_equal_or_both_nan(a, b):
    a, b = np.asarray(a), np.asarray(b)
    nan_a, nan_b = np.isnan(a), np.isnan(b)
    return np.array_equal(nan_a, nan_b) and np.array_equal(a.view(np.uint32)[~nan_a],
                                                           b.view(np.uint32)[~nan_b])


def _edge_floats():
    specials = [0.0, -0.0, 1.0, -1.0, 65504.0, 65505.0, 65519.99, 65520.0, 65536.0, 1e9,
                np.inf, -np.inf, 2.0 ** -14, 2.0 ** -15, 2.0 ** -24, 2.0 ** -25,
                2.0 ** -25 * 1.0000001, 3 * 2.0 ** -26, 2.0 ** -26, 1e-30, 1e-45, 3.4e38]
    # exact ties between consecutive halves: m * 2^(e-10) + half ulp
    ties = []
    for e in (-14, -3, 0, 5, 15):
        for m in (1024, 1025, 1026, 2047):
            ulp = 2.0 ** (e - 10)
            ties += [m * ulp + ulp / 2, -(m * ulp + ulp / 2)]
    # subnormal ties
    ties += [(k + 0.5) * 2.0 ** -24 for k in range(0, 8)]
    return np.array(specials + ties, np.float32)


def test_f16_widening_exhaustive(oracle):
    bits = np.arange(65536, dtype=np.uint16)
    ours = oracle.from_format(bits, "f16")
    ref = bits.view(np.float16).astype(np.float32)
    assert _bits_equal_or_both_nan(ours, ref)


def test_f16_narrowing_matches_numpy(oracle):
    rng = np.random.default_rng(16)
    rnd = rng.integers(0, 2 ** 32, 4_000_000, dtype=np.uint64).astype(np.uint32).view(np.float32)
    # dense sampling of the finite half range, where rounding matters most
    dense = (rng.uniform(-70000, 70000, 2_000_000) *
             np.exp2(rng.integers(-30, 1, 2_000_000))).astype(np.float32)
    for a in (rnd, dense, _edge_floats()):
        ours = oracle.to_format(a, "f16")
        with np.errstate(over="ignore"):
            ref = a.astype(np.float16)
        nan = np.isnan(a)
        assert np.array_equal(ours[~nan], ref.view(np.uint16)[~nan])
        assert np.isnan(ours[nan].view(np.float16)).all()


def test_bf16_matches_torch(oracle):
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(17)
    rnd = rng.integers(0, 2 ** 32, 4_000_000, dtype=np.uint64).astype(np.uint32).view(np.float32)
    for a in (rnd, _edge_floats()):
        ours = oracle.to_format(a, "bf16")
        ref = torch.from_numpy(a.copy()).to(torch.bfloat16).view(torch.int16).numpy()
        nan = np.isnan(a)
        assert np.array_equal(ours[~nan], ref.view(np.uint16)[~nan])
        back = oracle.from_format(ours, "bf16")
        assert np.isnan(back[nan]).all()
        # widening is exact: bf16 -> f32 -> bf16 is the identity
        assert np.array_equal(oracle.to_format(back[~nan], "bf16"), ours[~nan])


def _field(rng, shape, fmt, scale=1.0):
    return oracle_fmt(rng.uniform(-scale, scale, shape).astype(np.float32), fmt)


def oracle_fmt(a, fmt):
    import pyoracle
    return pyoracle.to_format(a, fmt)


@pytest.mark.parametrize("fmt", FMT16)
def test_step_is_rounded_float_step(oracle, fmt):
    rng = np.random.default_rng(3)
    shape = (20, 22, 26)
    P, PP = _field(rng, shape, fmt), _field(rng, shape, fmt)
    v2 = rng.uniform(0.0, 0.15, shape).astype(np.float32)
    coefs = oracle.stencil_coefs(8, 10.0, 11.0, 9.0)
    Pf, PPf = oracle.from_format(P, fmt), oracle.from_format(PP, fmt)
    assert oracle.step(Pf, PPf, v2, coefs, 8) == 0
    assert oracle.step(P, PP, v2, coefs, 8, fmt) == 0
    assert np.array_equal(PP, oracle.to_format(PPf, fmt))


@pytest.mark.parametrize("fmt", FMT16)
def test_upper_taper_source_receivers_upb_image_are_rounded(oracle, fmt):
    rng = np.random.default_rng(4)
    shape, b, order, nt = (24, 26, 28), 6, 8, 3
    g = oracle.taper_table(b, np.float32(0.3))
    P = _field(rng, shape, fmt)
    Pf = oracle.from_format(P, fmt)
    oracle.taper(P, g, b, True, fmt)
    oracle.taper(Pf, g, b, True)
    assert np.array_equal(P, oracle.to_format(Pf, fmt))
    # source: one rounding of (pp + s)
    PP = _field(rng, shape, fmt)
    PPf = oracle.from_format(PP, fmt)
    oracle.apply_source(PP, 7, 8, 9, np.float32(0.3712), fmt)
    oracle.apply_source(PPf, 7, 8, 9, np.float32(0.3712))
    assert np.array_equal(PP, oracle.to_format(PPf, fmt))
    # receivers: save widens, restore rounds
    geom = oracle.RcvGeom(1, 2, 3, 3, 4, 5, 4)
    rcv16 = np.zeros((5, 4, nt), np.float32)
    oracle.receivers(PP, rcv16, b, geom, nt, 1, False, fmt)
    rcvf = np.zeros((5, 4, nt), np.float32)
    oracle.receivers(oracle.from_format(PP, fmt), rcvf, b, geom, nt, 1, False)
    assert np.array_equal(rcv16, rcvf)
    rcv = rng.uniform(-1, 1, (5, 4, nt)).astype(np.float32)
    PPf = oracle.from_format(PP, fmt)
    oracle.receivers(PP, rcv, b, geom, nt, 2, True, fmt)
    oracle.receivers(PPf, rcv, b, geom, nt, 2, True)
    assert np.array_equal(PP, oracle.to_format(PPf, fmt))
    # UPB: an exact copy of the 16-bit layer, both ways
    upb = np.zeros((nt, shape[0], shape[1], order // 2), np.uint16)
    oracle.upb(PP, upb, b, order, 1, False, fmt)
    assert np.array_equal(upb[1], np.transpose(PP[:, :, b - 4:b], (0, 1, 2)))
    Q = np.zeros_like(PP)
    oracle.upb(Q, upb, b, order, 1, True, fmt)
    assert np.array_equal(Q[:, :, b - 4:b], PP[:, :, b - 4:b])
    # image: float accumulation of widened products
    PS, PR = _field(rng, shape, fmt), _field(rng, shape, fmt)
    img = rng.uniform(-1, 1, shape).astype(np.float32)
    imgf = img.copy()
    oracle.image(img, PS, PR, b, fmt)
    oracle.image(imgf, oracle.from_format(PS, fmt), oracle.from_format(PR, fmt), b)
    assert np.array_equal(img, imgf)


@pytest.mark.parametrize("fmt", FMT16)
def test_taper_all_rounds_per_pass_including_the_quirk(oracle, fmt):
    """Each region multiplies in float and rounds once at its store; the
    front-right z-edge reads the face cell (b, ny-b+iy, z) after the face
    pass has rounded it."""
    rng = np.random.default_rng(5)
    shape, b = (14, 15, 16), 4
    nx, ny, nz = shape
    g = oracle.taper_table(b, np.float32(0.2))
    P = _field(rng, shape, fmt)
    P0 = oracle.from_format(P, fmt)
    Pf = P0.copy()
    oracle.taper(P, g, b, False, fmt)
    oracle.taper(Pf, g, b, False)
    quirk = np.zeros(shape, bool)
    quirk[:b, ny - b:, b:nz - b] = True
    assert np.array_equal(P[~quirk], oracle.to_format(Pf, fmt)[~quirk])
    # quirk cells by hand: face pass on x (ix = b is interior in x, so the
    # y-face pass tapered (b, ny-b+iy, z) by g[iy] and rounded it)
    rnd = lambda v: oracle.from_format(oracle.to_format(np.float32(v).reshape(1), fmt), fmt)[0]
    for ix in range(b):
        for iy in range(b):
            for iz in range(b, nz - b):
                face = rnd(np.float32(P0[b, ny - b + iy, iz]) * g[iy])
                want = rnd(np.float32(np.float32(face * g[b - 1 - ix]) * g[iy]))
                got = oracle.from_format(P[ix:ix + 1, ny - b + iy, iz], fmt)[0]
                assert got == want


@pytest.mark.parametrize("fmt, bound", [("f16", 1e-2), ("bf16", 1e-1)])
def test_modeling_error_vs_f64(oracle, fmt, bound):
    """Traces of a 40^3 3-layer modeling run (300 steps) with 16-bit
    wavefields, against the FP64 oracle.  Measured relative L2: f16 4.6e-3,
    bf16 4.9e-2 (11- and 8-bit significands rounded every step), FP32
    4.7e-6; the bounds leave about 2x headroom."""
    from rtm3d_demo_amd.workloads import cfl_dt, params, three_layer
    p = params("l3_40", 40, 40, 40, 10.0, 10, 300, cfl_dt(3500.0, 10.0), 15.0, (20, 20, 2),
               (0, 0, 0, 2, 2, 20, 20))
    vel = three_layer(40, 40, 40, 13, 26)
    ref = oracle.pipeline_modeling(p, vel, np.float64)
    f32 = oracle.pipeline_modeling(p, vel, np.float32)
    got = oracle.pipeline_modeling(p, vel, np.float32, fmt)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    err32 = np.linalg.norm(f32 - ref) / np.linalg.norm(ref)
    assert err32 < 1e-5
    assert err < bound, err
