"""Pins of the oracle that do not depend on its own formulas: exact
rationals and moments of the taps, polynomial exactness of the stencil,
axis-resolved impulse responses, closed forms of the Ricker wavelet and the
Gaussian taper, index encodings of receivers / upper border / imaging, and
time-reversibility of the propagator."""
from fractions import Fraction

import numpy as np
import pytest

TAPS = {
    2: [Fraction(1), Fraction(-2), Fraction(1)],
    4: [Fraction(-1, 12), Fraction(4, 3), Fraction(-5, 2), Fraction(4, 3), Fraction(-1, 12)],
    6: [Fraction(1, 90), Fraction(-3, 20), Fraction(3, 2), Fraction(-49, 18), Fraction(3, 2),
        Fraction(-3, 20), Fraction(1, 90)],
    8: [Fraction(-1, 560), Fraction(8, 315), Fraction(-1, 5), Fraction(8, 5), Fraction(-205, 72),
        Fraction(8, 5), Fraction(-1, 5), Fraction(8, 315), Fraction(-1, 560)],
}


@pytest.mark.parametrize("order", [2, 4, 6, 8])
def test_taps_are_the_textbook_rationals(oracle, order):
    taps = oracle.laplacian_taps(order)
    R = order // 2
    exact = TAPS[order]
    assert np.array_equal(taps, np.array([float(f) for f in exact], np.float32))
    # central-difference moments: sum 0, second moment 2, higher even moments 0
    assert sum(exact) == 0
    assert sum(c * (k - R) ** 2 for k, c in enumerate(exact)) == 2
    for m in range(2, R + 1):
        assert sum(c * (k - R) ** (2 * m) for k, c in enumerate(exact)) == 0


def test_stencil_plane_scaling(oracle):
    c = oracle.stencil_coefs(8, 10.0, 12.5, 8.0).reshape(3, 9)
    taps = oracle.laplacian_taps(8)
    for k, d in enumerate((10.0, 12.5, 8.0)):
        inv = np.float32((1.0 / np.float64(np.float32(d))) ** 2)
        assert np.array_equal(c[k], (inv * taps).astype(np.float32))


def _grid(shape, fn, dtype):
    x, y, z = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij")
    return np.ascontiguousarray(fn(x, y, z), dtype)


@pytest.mark.parametrize("order", [2, 4, 8])
def test_step_exact_on_polynomials(oracle, order):
    """The order-2R stencil differentiates polynomials of degree <= 2R+1
    exactly: with PP = 0 and V = 1, PP_out - 2P = lap(P)."""
    d = (9.0, 11.0, 7.0)
    shape = (17, 19, 21)
    deg = 2 * (order // 2) + 1
    a = min(deg, 4)
    b = min(deg, 3)
    P = _grid(shape, lambda x, y, z: (x * d[0]) ** a + (y * d[1]) ** b * (z * d[2]) ** 2
              + (z * d[2]) ** min(deg, 5), np.float64)
    lap = _grid(shape, lambda x, y, z: a * (a - 1) * (x * d[0]) ** (a - 2)
                + b * (b - 1) * (y * d[1]) ** (b - 2) * (z * d[2]) ** 2
                + 2 * (y * d[1]) ** b
                + min(deg, 5) * (min(deg, 5) - 1) * (z * d[2]) ** (min(deg, 5) - 2), np.float64)
    coefs = oracle.stencil_coefs(order, *d, dtype=np.float64)
    PP = np.zeros_like(P)
    V = np.ones_like(P)
    assert oracle.step(P, PP, V, coefs, order) == 0
    R = order // 2
    got = (PP - 2 * P)[R:-R, R:-R, R:-R]
    want = lap[R:-R, R:-R, R:-R]
    assert np.max(np.abs(got - want)) <= 1e-9 * np.max(np.abs(want))
    # cells outside [R, n-R) are untouched
    assert np.all(PP[:R] == 0) and np.all(PP[:, -R:] == 0) and np.all(PP[:, :, :R] == 0)


def test_step_impulse_reveals_each_axis_taps(oracle):
    """An impulse's response along axis d is v*(p0*c_d[R-k]) exactly."""
    order, R = 8, 4
    coefs = oracle.stencil_coefs(order, 10.0, 12.0, 8.0)
    c = coefs.reshape(3, 9)
    shape = (21, 23, 25)
    cx, cy, cz = 10, 11, 12
    P = np.zeros(shape, np.float32)
    P[cx, cy, cz] = np.float32(3.0)
    PP = np.zeros(shape, np.float32)
    V = np.full(shape, np.float32(0.7), np.float32)
    oracle.step(P, PP, V, coefs, order)
    p0, v = np.float32(3.0), np.float32(0.7)
    for k in range(1, R + 1):
        for s in (-1, 1):
            assert PP[cx + s * k, cy, cz] == v * (p0 * c[0][R - s * k])
            assert PP[cx, cy + s * k, cz] == v * (p0 * c[1][R - s * k])
            assert PP[cx, cy, cz + s * k] == v * (p0 * c[2][R - s * k])
    nz = np.argwhere(PP != 0)
    assert len(nz) == 6 * R + 1  # the star and its centre, nothing else


def test_step_f32_matches_vectorised_float32(oracle):
    """Same arithmetic evaluated by numpy float32 array operations in the
    reference's order: bit-exact agreement on a random non-cubic grid."""
    rng = np.random.default_rng(3)
    shape = (13, 17, 29)
    P = rng.standard_normal(shape).astype(np.float32)
    PP = rng.standard_normal(shape).astype(np.float32)
    V = rng.uniform(0.01, 0.2, shape).astype(np.float32)
    coefs = oracle.stencil_coefs(8, 10.0, 12.0, 8.0)
    c = coefs.reshape(3, 9)
    want = PP.copy()
    R = 4
    core = (slice(R, -R),) * 3
    ax = np.zeros(P[core].shape, np.float32)
    ay, az = ax.copy(), ax.copy()
    nx, ny, nz = shape
    for io in range(9):
        o = io - R
        ax = ax + P[R + o:nx - R + o, R:-R, R:-R] * c[0][io]
        ay = ay + P[R:-R, R + o:ny - R + o, R:-R] * c[1][io]
        az = az + P[R:-R, R:-R, R + o:nz - R + o] * c[2][io]
    lapl = (ax + ay) + az
    want[core] = (np.float32(2) * P[core] - PP[core]) + V[core] * lapl
    got = PP.copy()
    oracle.step(P, got, V, coefs, 8)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_step_reports_nan(oracle):
    shape = (12, 12, 12)
    P = np.zeros(shape, np.float32)
    P[6, 6, 6] = np.nan
    PP = np.zeros(shape, np.float32)
    assert oracle.step(P, PP, np.ones(shape, np.float32), oracle.stencil_coefs(8, 1, 1, 1), 8) == -1


def test_ricker_closed_form(oracle):
    nt, dt, fp = 400, np.float32(1e-3), np.float32(15.0)
    s = oracle.ricker(nt, dt, fp).astype(np.float64)
    t = np.arange(nt) * float(dt) - 1.0 / float(fp)
    x = np.pi * float(fp) * t
    exact = (1 - 2 * x * x) * np.exp(-x * x)
    assert np.max(np.abs(s - exact)) < 2e-6
    # peak at t = 1/fpeak
    assert abs(s.max() - 1.0) < 1e-3


def test_taper_closed_form(oracle):
    for blen, f in ((40, 0.9), (10, 0.5), (32, 0.98)):
        g = oracle.taper_table(blen, np.float32(f)).astype(np.float64)
        dfrac = np.sqrt(-np.log(f)) / blen
        exact = np.exp(-(dfrac * np.arange(blen)) ** 2)
        assert g[0] == 1.0
        assert np.max(np.abs(g - exact)) < 1e-6


def _constant_taper(oracle, shape, blen, upper):
    g = oracle.taper_table(blen, np.float32(0.6))
    P = np.ones(shape, np.float32)
    oracle.taper(P, g, blen, upper)
    return P, g.astype(np.float64)


def test_taper_all_on_constant_field(oracle):
    """All-border taper of a constant field equals the product of the axis
    factors, except the front-right edge, which takes the face-tapered value
    of (blen, y, z) and applies the y factor a second time."""
    shape, b = (14, 16, 18), 4
    P, g = _constant_taper(oracle, shape, b, False)

    def f(i, n):
        return g[b - 1 - i] if i < b else (g[i - (n - b)] if i >= n - b else 1.0)

    for x in range(shape[0]):
        for y in range(shape[1]):
            for z in range(shape[2]):
                want = f(x, shape[0]) * f(y, shape[1]) * f(z, shape[2])
                front_right = x < b and y >= shape[1] - b and b <= z < shape[2] - b
                if front_right:
                    want = f(y, shape[1]) ** 2 * f(x, shape[0])
                assert abs(P[x, y, z] - want) <= 3e-7, (x, y, z)


def test_taper_upper_on_constant_field(oracle):
    shape, b = (14, 16, 18), 4
    P, g = _constant_taper(oracle, shape, b, True)

    def f(i, n):
        return g[b - 1 - i] if i < b else (g[i - (n - b)] if i >= n - b else 1.0)

    for x in range(shape[0]):
        for y in range(shape[1]):
            for z in range(shape[2]):
                want = f(x, shape[0]) * f(y, shape[1]) * f(z, shape[2]) if z < b else 1.0
                assert abs(P[x, y, z] - want) <= 3e-7, (x, y, z)


def _coded(shape):
    x, y, z = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    return (x * 10000 + y * 100 + z).astype(np.float32)


def test_receivers_index_encoding(oracle):
    shape, b = (30, 34, 26), 5
    g = oracle.RcvGeom(1, 2, 3, 2, 3, 9, 8)  # start x,y, depth, dist x,y, count x,y
    P = _coded(shape)
    rcv = np.zeros((9, 8, 7), np.float32)
    oracle.receivers(P, rcv, b, g, 7, 4, False)
    for ix in range(9):
        for iy in range(8):
            assert rcv[ix, iy, 4] == (1 + 2 * ix + b) * 10000 + (2 + 3 * iy + b) * 100 + 3 + b
    Q = np.zeros(shape, np.float32)
    oracle.receivers(Q, rcv, b, g, 7, 4, True)
    mask = Q != 0
    assert mask.sum() == 9 * 8 and np.array_equal(Q[mask], P[mask])


def test_upper_border_index_encoding(oracle):
    shape, b, order = (11, 13, 20), 6, 8
    P = _coded(shape)
    upb = np.zeros((3, 11, 13, 4), np.float32)
    oracle.upb(P, upb, b, order, 2, False)
    assert np.array_equal(upb[2], P[:, :, b - 4:b])
    Q = np.zeros(shape, np.float32)
    oracle.upb(Q, upb, b, order, 2, True)
    assert np.array_equal(Q[:, :, b - 4:b], P[:, :, b - 4:b])
    assert np.count_nonzero(Q[:, :, :b - 4]) == 0 and np.count_nonzero(Q[:, :, b:]) == 0


def test_image_range_and_finish(oracle):
    shape, b = (12, 14, 16), 3
    img = np.zeros(shape, np.float32)
    ones = np.ones(shape, np.float32)
    oracle.image(img, ones, ones * 2, b)
    assert np.all(img[:, :, b:-b] == 2) and np.all(img[:, :, :b] == 0) and np.all(img[:, :, -b:] == 0)
    # Laplacian filter of a quadratic, cropped: 2/dx^2 + 2/dy^2 + 2/dz^2
    d = (2.0, 4.0, 0.5)
    q = _grid(shape, lambda x, y, z: (x * d[0]) ** 2 + (y * d[1]) ** 2 + (z * d[2]) ** 2,
              np.float64)
    out = oracle.image_finish(q, b, *d)
    assert out.shape == (12 - 2 * b, 14 - 2 * b, 16 - 2 * b)
    assert np.allclose(out, 6.0, rtol=0, atol=1e-9)


def test_propagator_time_reversible_f64(oracle):
    """With no taper (factor 1) and a source that is silent after step K,
    stepping back from the last two fields with the upper layer restored
    reproduces the forward fields (f64, to rounding)."""
    shape, b, order = (30, 32, 34), 8, 8
    nt, K = 120, 30
    rng = np.random.default_rng(0)
    V = rng.uniform(0.02, 0.12, shape)
    coefs = oracle.stencil_coefs(order, 1.0, 1.0, 1.0, dtype=np.float64)
    g = np.ones(b)
    src = oracle.ricker(nt, 0.004, 20.0, np.float64)
    src[K:] = 0
    P, PP = np.zeros(shape), np.zeros(shape)
    upb = np.zeros((nt, shape[0], shape[1], 4))
    hist = {}
    for it in range(nt):
        P, PP = PP, P
        oracle.taper(P, g, b, True)
        oracle.taper(PP, g, b, True)
        oracle.step(P, PP, V, coefs, order)
        oracle.apply_source(PP, 15, 16, b + 3, src[it])
        oracle.upb(PP, upb, b, order, it, False)
        hist[it] = PP.copy()
    Ps, PPs = P.copy(), PP.copy()  # p(nt-2), p(nt-1)
    for it in range(nt - 3, K - 1, -1):
        oracle.step(Ps, PPs, V, coefs, order)
        oracle.upb(PPs, upb, b, order, it, True)
        Ps, PPs = PPs, Ps
        err = np.abs(Ps - hist[it])[:, :, b - 4:].max() / np.abs(hist[it]).max()
        assert err < 1e-9, (it, err)
