"""CPU oracle of the snapshot autoencoder -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU leg may import this
module, as the checker of the device kernels; the product never calls it.

The network has no reference counterpart (the reference keeps the source
field by HBC reconstruction, rtm-3d/src/kernel/RTMFiniteDifferencesKernel_HBC.cpp,
or by exact snapshots).  This is a float64 numpy restatement of the device
network (rtm3d_demo_amd/csrc/rtm_ae.hip):

  brick order  element e = ((a*64 + m)*8 + l) of a 32^3 brick is the cell
               (8 ax + 2 mx + lx, 8 ay + 2 my + ly, 8 az + 2 mz + lz),
               a = (ax*4+ay)*4+az, m = (mx*4+my)*4+mz, l = (lx*2+ly)*2+lz
  L1  H1 = tanh(X[4096n x 8] W1 + b1)          2^3 patches, 1 -> 8
  L2  Z  = H1[64n x 512] W2 + b2                4^3 patches, 8 -> 32 (code)
  L3  H3 = tanh(Z W3 + b3)                      transposed L2, 32 -> 8
  L4  Y  = H3[4096n x 8] W4 + b4                transposed L1, 8 -> 1
  loss = 0.5 * sum (Y - X)^2

Pinned by central finite differences of the loss (tests/test_ae_oracle.py)
and by the geometry of the brick order (patch rows are aligned cubes).
"""
import numpy as np

LAYERS = [(8, 8, True, 4096), (512, 32, False, 64), (32, 512, True, 64), (8, 8, False, 4096)]
BRICK = 32768
CODES = 2048
NPARAMS = sum(k * n + n for k, n, _, _ in LAYERS)


def unpack(params):
    """[(W[K][N], b[N])] views of the flat parameter vector."""
    out, o = [], 0
    for k, n, _, _ in LAYERS:
        W = params[o:o + k * n].reshape(k, n)
        o += k * n
        b = params[o:o + n]
        o += n
        out.append((W, b))
    return out


def brick_cells():
    """(x, y, z) of every element of a brick, in brick order."""
    e = np.arange(BRICK)
    l, m, a = e & 7, (e >> 3) & 63, e >> 9
    x = 8 * (a >> 4) + 2 * (m >> 4) + (l >> 2)
    y = 8 * ((a >> 2) & 3) + 2 * ((m >> 2) & 3) + ((l >> 1) & 1)
    z = 8 * (a & 3) + 2 * (m & 3) + (l & 1)
    return x, y, z


def region_bricks(region):
    x0, y0, z0, x1, y1, z1 = region
    return (-(-(x1 - x0) // 32), -(-(y1 - y0) // 32), -(-(z1 - z0) // 32))


def bricks(field, region):
    """FP16 bricks (brick order, normalised by the largest magnitude) and
    FP32 scales of a region of an FP32 field."""
    x0, y0, z0, x1, y1, z1 = region
    nbx, nby, nbz = region_bricks(region)
    cx, cy, cz = brick_cells()
    out = np.zeros((nbx * nby * nbz, BRICK), np.float16)
    scales = np.zeros(nbx * nby * nbz, np.float32)
    k = 0
    for bx in range(nbx):
        for by in range(nby):
            for bz in range(nbz):
                x, y, z = x0 + 32 * bx + cx, y0 + 32 * by + cy, z0 + 32 * bz + cz
                inside = (x < x1) & (y < y1) & (z < z1)
                v = np.zeros(BRICK, np.float32)
                v[inside] = field[x[inside], y[inside], z[inside]]
                s = np.float32(np.abs(v).max())
                if s < np.finfo(np.float32).tiny:  # zero or subnormal: a zero brick
                    s = np.float32(0)
                scales[k] = s
                if s > 0:
                    out[k] = (v * (np.float32(1.0) / s)).astype(np.float16)
                k += 1
    return out, scales


def unbricks(Y, scales, field, region):
    """Writes Y[k] * scales[k] into the region's cells of field."""
    x0, y0, z0, x1, y1, z1 = region
    nbx, nby, nbz = region_bricks(region)
    cx, cy, cz = brick_cells()
    k = 0
    for bx in range(nbx):
        for by in range(nby):
            for bz in range(nbz):
                x, y, z = x0 + 32 * bx + cx, y0 + 32 * by + cy, z0 + 32 * bz + cz
                inside = (x < x1) & (y < y1) & (z < z1)
                field[x[inside], y[inside], z[inside]] = (Y[k][inside] * scales[k]).astype(
                    field.dtype)
                k += 1


def forward(params, X):
    """X: [n, 32768] (any float type) -> (Y [n, 32768], cache), float64."""
    (W1, b1), (W2, b2), (W3, b3), (W4, b4) = [(w.astype(np.float64), b.astype(np.float64))
                                              for w, b in unpack(params)]
    n = X.shape[0]
    X1 = X.astype(np.float64).reshape(n * 4096, 8)
    H1 = np.tanh(X1 @ W1 + b1)
    Z = H1.reshape(n * 64, 512) @ W2 + b2
    H3 = np.tanh(Z @ W3 + b3)
    Y = H3.reshape(n * 4096, 8) @ W4 + b4
    return Y.reshape(n, BRICK), (X1, H1, Z, H3)


def encode(params, X):
    (W1, b1), (W2, b2) = [(w.astype(np.float64), b.astype(np.float64))
                          for w, b in unpack(params)[:2]]
    n = X.shape[0]
    H1 = np.tanh(X.astype(np.float64).reshape(n * 4096, 8) @ W1 + b1)
    return (H1.reshape(n * 64, 512) @ W2 + b2).reshape(n, CODES)


def decode(params, Z):
    (W3, b3), (W4, b4) = [(w.astype(np.float64), b.astype(np.float64))
                          for w, b in unpack(params)[2:]]
    n = Z.shape[0]
    H3 = np.tanh(Z.astype(np.float64).reshape(n * 64, 32) @ W3 + b3)
    return (H3.reshape(n * 4096, 8) @ W4 + b4).reshape(n, BRICK)


def loss_and_grad(params, X):
    """0.5 * sum (Y - X)^2 and its gradient (flat, parameter layout)."""
    (W1, _), (W2, _), (W3, _), (W4, _) = [(w.astype(np.float64), b) for w, b in unpack(params)]
    n = X.shape[0]
    Y, (X1, H1, Z, H3) = forward(params, X)
    dY = (Y - X.astype(np.float64)).reshape(n * 4096, 8)
    loss = 0.5 * float((dY * dY).sum())
    H3v = H3.reshape(n * 4096, 8)
    gW4, gb4 = H3v.T @ dY, dY.sum(0)
    dA3 = (dY @ W4.T).reshape(n * 64, 512) * (1.0 - H3 * H3)
    gW3, gb3 = Z.T @ dA3, dA3.sum(0)
    dZ = dA3 @ W3.T
    H1v = H1.reshape(n * 64, 512)
    gW2, gb2 = H1v.T @ dZ, dZ.sum(0)
    dA1 = (dZ @ W2.T).reshape(n * 4096, 8) * (1.0 - H1 * H1)
    gW1, gb1 = X1.T @ dA1, dA1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3, gW4.ravel(), gb4])
    return loss, g
