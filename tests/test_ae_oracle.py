"""Pins the autoencoder oracle (oracle/ae_oracle.py) without restating it:
central finite differences of its loss against its analytic gradient, the
geometry of the brick order (patch rows are aligned cubes), and the
brick cut / paste round trip."""
import numpy as np

import ae_oracle as ao


def _params(rng, scale=1.0):
    p = np.zeros(ao.NPARAMS)
    o = 0
    for k, n, _, _ in ao.LAYERS:
        p[o:o + k * n] = rng.uniform(-1, 1, k * n) * scale * np.sqrt(6.0 / (k + n))
        o += k * n
        p[o:o + n] = rng.uniform(-0.1, 0.1, n)
        o += n
    return p


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(11)
    p = _params(rng)
    X = rng.uniform(-1, 1, (2, ao.BRICK))
    _, g = ao.loss_and_grad(p, X)
    # a few entries of every layer's weights and biases
    o, picks = 0, []
    for k, n, _, _ in ao.LAYERS:
        picks += list(o + rng.choice(k * n, 6, replace=False))
        o += k * n
        picks += list(o + rng.choice(n, 3, replace=False))
        o += n
    h = 1e-5
    for i in picks:
        q = p.copy()
        q[i] += h
        lp, _ = ao.loss_and_grad(q, X)
        q[i] -= 2 * h
        lm, _ = ao.loss_and_grad(q, X)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(fd)), (i, fd, g[i])


def test_brick_order_is_nested_cubes():
    x, y, z = ao.brick_cells()
    lin = (x * 32 + y) * 32 + z
    assert np.array_equal(np.sort(lin), np.arange(ao.BRICK))  # a permutation of 32^3
    for size, span in ((8, 2), (512, 8)):  # L1 rows: 2^3 cubes; L2 rows: 8^3 cubes
        for arr in (x, y, z):
            r = arr.reshape(-1, size)
            assert (r.min(1) % span == 0).all() and (r.max(1) - r.min(1) == span - 1).all()


def test_bricks_round_trip_on_a_partial_region():
    rng = np.random.default_rng(12)
    f = rng.uniform(-3, 3, (40, 37, 50)).astype(np.float32)
    region = (2, 1, 6, 40, 37, 45)  # partial bricks on every axis
    B, s = ao.bricks(f, region)
    assert B.shape == (2 * 2 * 2, ao.BRICK)
    out = np.zeros_like(f)
    ao.unbricks(B.astype(np.float32), s, out, region)
    x0, y0, z0, x1, y1, z1 = region
    inside = np.zeros(f.shape, bool)
    inside[x0:x1, y0:y1, z0:z1] = True
    assert (out[~inside] == 0).all()
    err = np.abs(out - f)[inside].max()
    assert err <= 3.0 * 2.0 ** -11 * 1.01  # one FP16 rounding of values <= 1, times the scale
